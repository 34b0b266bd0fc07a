"""Tiles holding solid cells swept as general tiles ("sweep") or with their
fluid cells in the fix-up list ("cells", the one-kernel step of the float
forms).

FP64 in "cells" mode against the oracle (bar 1e-12, expected bitwise); every
storage form bitwise equal between the two modes, since each cell's
arithmetic and its stochastic-rounding bits do not depend on the path that
sweeps it; and a 16-bit run reproduced bit for bit.
"""
import numpy as np
import pytest

from oracle import d2q9_oracle as O
from pylabolt_amd import cases
from tests.helpers import max_rel, oracle_case

pytestmark = pytest.mark.gpu

CYLINDERS = {
    "cylinder_zg_pressure": lambda: cases.cylinder(40, 20, diameter=8, xc=12),
    "cylinder_wide": lambda: cases.cylinder(44, 130, diameter=10, xc=14),
    "cylinder_two_blocks": lambda: cases.cylinder(64, 300, diameter=24,
                                                  xc=20),
}


def _solver(sim, storage, mode, graph_steps=4):
    from pylabolt_amd.solvers.fluidLB import Solver
    s = Solver(simulation=sim, storage=storage, graph_steps=graph_steps,
               verbose=False, solid_tiles=mode)
    s.set_backend()
    return s


@pytest.mark.parametrize("name", sorted(CYLINDERS))
def test_cells_mode_fp64_exact(name):
    sim = CYLINDERS[name]()
    case = oracle_case(sim)
    O.initialize_pop(case)
    s = _solver(sim, "f64", "cells")
    t = s.engine.tables
    assert t.solid_tiles == "cells" and t.n_general == 0
    assert (t.tile_class != 0).any()
    for chunk in (1, 6, 13):
        s.advance(chunk)
        O.run(case, chunk)
        fl = case.fluid
        rho, vel = s.engine.moments_host()
        errs = {"pop": max_rel(s.engine.pop()[fl], case.pop[fl]),
                "pop_new": max_rel(s.engine.pop_new()[fl], case.pop_new[fl]),
                "density": max_rel(rho[fl], case.density[fl]),
                "velocity": max_rel(vel[fl], case.velocity[fl])}
        print(f"[cells] {name} step {case.time_step}: {errs}", flush=True)
        for k, v in errs.items():
            assert v <= 1e-12, f"{name} {k} {v}"


@pytest.mark.parametrize("storage", ["f32", "f16dev", "bf16dev"])
@pytest.mark.parametrize("name", ["cylinder_wide", "cylinder_two_blocks"])
def test_modes_bitwise_equal(name, storage):
    sweep = _solver(CYLINDERS[name](), storage, "sweep")
    cells = _solver(CYLINDERS[name](), storage, "cells")
    assert sweep.engine.tables.n_general > 0
    assert cells.engine.tables.n_general == 0
    done = 0
    for chunk in (1, 9, 10):
        sweep.advance(chunk)
        cells.advance(chunk)
        done += chunk
        for what in ("pop", "pop_new"):
            a = getattr(sweep.engine, what)()
            b = getattr(cells.engine, what)()
            assert np.array_equal(a, b), f"{name} {storage} {what} {done}"
        ra, va = sweep.engine.moments_host()
        rb, vb = cells.engine.moments_host()
        assert np.array_equal(ra, rb) and np.array_equal(va, vb)
        print(f"[cells] {name} {storage} bitwise equal after {done} steps",
              flush=True)


@pytest.mark.parametrize("storage", ["f16dev", "bf16dev"])
def test_stochastic_rounding_reproducible(storage):
    runs = []
    for _ in range(2):
        s = _solver(CYLINDERS["cylinder_two_blocks"](), storage, None)
        s.advance(17)
        runs.append(s.engine.pop())
    assert np.array_equal(runs[0], runs[1])
