"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py).

CPU: the oracle still reproduces them bit for bit (regression of the
checker).  GPU: the HIP engine matches them to the FP64 bar
(max relative error <= 1e-12)."""
import os

import numpy as np
import pytest

from oracle import d2q9_oracle as O
from tests.golden.make_golden import golden_cases
from tests.helpers import max_rel, oracle_case

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NAMES = sorted(golden_cases())


def _load(name):
    return np.load(os.path.join(HERE, name + ".npz"))


@pytest.mark.parametrize("name", NAMES)
def test_oracle_reproduces_golden(name):
    sim, steps = golden_cases()[name]
    g = _load(name)
    assert int(g["steps"]) == steps
    case = oracle_case(sim)
    O.initialize_pop(case)
    O.run(case, steps)
    for key in ("pop", "pop_new", "density", "velocity"):
        assert np.array_equal(getattr(case, key), g[key]), key


@pytest.mark.gpu
@pytest.mark.parametrize("name", NAMES)
def test_engine_matches_golden(name):
    from pylabolt_amd.solvers.fluidLB import Solver
    sim, steps = golden_cases()[name]
    g = _load(name)
    s = Solver(simulation=sim, verbose=False, graph_steps=4)
    s.set_backend()
    s.advance(steps)
    eng = s.engine
    fl = g["fluid"]
    rho, vel = eng.moments_host()
    errs = [max_rel(eng.pop()[fl], g["pop"][fl]),
            max_rel(eng.pop_new()[fl], g["pop_new"][fl]),
            max_rel(rho[fl], g["density"][fl]),
            max_rel(vel[fl], g["velocity"][fl])]
    print(f"[golden] {name}: {errs}", flush=True)
    assert max(errs) <= 1e-12
