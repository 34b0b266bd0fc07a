"""Pairs of steps as one sweep over the deep interior (d2q9_fused_step2)
against the oracle (FP64) and against single steps (FP64 and FP32).

Bar: max relative error 1e-12 against the oracle (expected bitwise), and
bitwise equality with the single-step path, after odd and even step counts,
with and without graph replay, with the moments stored at the last step.
The lattices are wide enough (ny >= 2 blocks + 2) for deep tiles of both
block widths (64 cells in FP64, 128 in FP32).
"""
import numpy as np
import pytest

from oracle import d2q9_oracle as O
from pylabolt_amd import cases
from tests.helpers import max_rel, oracle_case

pytestmark = pytest.mark.gpu


def _rough(sim):
    sim.initial_fields_dict["default"]["fluid"]["velocity"] = {
        "type": "func",
        "func": lambda i, j: (0.03 * np.sin(1.3 * i + 2.1 * j),
                              0.03 * np.cos(0.7 * i - 1.9 * j))}
    return sim


def _forced_channel(ny=300):
    sim = cases.poiseuille_channel(80, ny, model="MRT",
                                   forcing="guo_second_order")
    sim.forcing_dict["gravity"] = [2e-5, 3e-6]
    return sim


def _periodic_box(nx=60, ny=300):
    """Periodic in x and y, Guo-forced: deep tiles take the periodic rows,
    the periodic columns stay in the rim."""
    box = cases.poiseuille_channel(nx, ny, model="BGK", forcing="guo_linear")
    box.boundary_dict = {
        "options": {},
        "x": {"wall": False,
              "segments": [[[0, 0], [0, ny - 1]], [[nx - 1, 0], [nx - 1, ny - 1]]],
              "fluid": {"type": "periodic"}},
        "y": {"wall": False,
              "segments": [[[0, 0], [nx - 1, 0]], [[0, ny - 1], [nx - 1, ny - 1]]],
              "fluid": {"type": "periodic"}},
    }
    box.forcing_dict["gravity"] = [1e-5, 2e-5]
    return box


CASES = {
    "cylinder": lambda: cases.cylinder(120, 300, diameter=20, xc=40),
    "cavity": lambda: cases.lid_driven_cavity(100, 300),
    "channel_mrt_guo2": _forced_channel,
    "cylinder_mrt": lambda: cases.cylinder(90, 260, diameter=16, xc=30,
                                           model="MRT"),
    # ny a multiple of both block widths: deep tiles reach both walls
    # (zero-gradient for the cylinder, bounce-back for the channel)
    "cylinder_walls_both": lambda: cases.cylinder(100, 256, diameter=16,
                                                  xc=50),
    "channel_walls_both": lambda: _forced_channel(256),
    # periodic rows read from their images and the last row of tiles moved
    # up to end at the last row (80 = 6 x 12 + 8): every cell deep, no rim;
    # and rows a multiple of the tile height
    "channel_rows_multiple": lambda: cases.poiseuille_channel(
        72, 256, model="BGK", forcing="guo_linear"),
    "periodic_box": _periodic_box,
}


def _solver(sim, two_step, graph_steps=8, storage="f64"):
    from pylabolt_amd.solvers.fluidLB import Solver
    s = Solver(simulation=sim, storage=storage, graph_steps=graph_steps,
               verbose=False, two_step=two_step)
    s.set_backend()
    return s


@pytest.mark.parametrize("name", sorted(CASES))
def test_two_step_matches_oracle(name):
    sim = _rough(CASES[name]())
    case = oracle_case(sim)
    O.initialize_pop(case)
    s = _solver(sim, True)
    eng = s.engine
    assert eng.two_step and eng.n_deep > 0
    if name.endswith("walls_both"):
        walls = eng._deep["walls"].cpu().numpy()[:eng.n_deep]
        assert (walls & 3).any() and (walls >> 2).any()
    if name in ("channel_walls_both", "channel_rows_multiple"):
        assert eng.deep_fraction == 1.0
    print(f"[two-step] {name}: {eng.n_deep} deep tiles, "
          f"{eng.deep_fraction:.0%} of the cells", flush=True)
    done = 0
    for chunk in (3, 2, 21):
        s.advance(chunk)
        O.run(case, chunk)
        done += chunk
        fl = case.fluid
        rho, vel = eng.moments_host()
        errs = {"pop": max_rel(eng.pop()[fl], case.pop[fl]),
                "pop_new": max_rel(eng.pop_new()[fl], case.pop_new[fl]),
                "density": max_rel(rho[fl], case.density[fl]),
                "velocity": max_rel(vel[fl], case.velocity[fl])}
        print(f"[two-step] {name} step {done}: {errs}", flush=True)
        for k, v in errs.items():
            assert v <= 1e-12, f"{name} step {done}: {k} {v}"


@pytest.mark.parametrize("storage", ["f64", "f32"])
@pytest.mark.parametrize("name", ["cylinder", "cavity", "channel_mrt_guo2",
                                  "cylinder_walls_both",
                                  "channel_walls_both"])
def test_two_step_bitwise_equals_single_steps(name, storage):
    one = _solver(CASES[name](), False, storage=storage)
    two = _solver(CASES[name](), True, storage=storage)
    assert two.engine.two_step and not one.engine.two_step
    done = 0
    for chunk, graph in ((1, True), (2, False), (37, True), (18, False)):
        one.engine.advance(chunk, use_graph=graph)
        two.engine.advance(chunk, use_graph=graph)
        done += chunk
        assert np.array_equal(one.engine.pop(), two.engine.pop()), done
        ra, ua = one.engine.moments_host()
        rb, ub = two.engine.moments_host()
        assert np.array_equal(ra, rb) and np.array_equal(ua, ub), done
        print(f"[two-step] {name} {storage} bitwise equal after {done} "
              f"steps", flush=True)


def test_two_step_off_without_deep_tiles():
    s = _solver(cases.lid_driven_cavity(20, 30), True)
    assert not s.engine.two_step
    s.advance(5)


ABS_TOL = {"f32": 2e-6, "f16dev": 1e-4, "bf16dev": 1e-3}


@pytest.mark.parametrize("storage", sorted(ABS_TOL))
@pytest.mark.parametrize("name", ["cylinder", "cavity", "channel_mrt_guo2"])
def test_two_step_reduced_precision_close_to_oracle(name, storage):
    """The float forms through the two-step sweep against the FP64 oracle
    after 20 steps, at the tolerances of the single-step precision tests
    (16-bit storage rounds only the stored steps here, so it is at least as
    close)."""
    sim = _rough(CASES[name]())
    case = oracle_case(sim)
    O.initialize_pop(case)
    s = _solver(sim, True, storage=storage)
    assert s.engine.two_step
    s.advance(20)
    O.run(case, 20)
    fl = case.fluid
    err_pop = float(np.abs(s.engine.pop()[fl] - case.pop[fl]).max())
    err_new = float(np.abs(s.engine.pop_new()[fl] - case.pop_new[fl]).max())
    _, vel = s.engine.moments_host()
    err_u = float(np.abs(vel[fl] - case.velocity[fl]).max())
    print(f"[two-step] {name} {storage}: pop {err_pop:.2e} pop_new "
          f"{err_new:.2e} u {err_u:.2e}", flush=True)
    tol = ABS_TOL[storage]
    assert err_pop <= tol and err_new <= tol and err_u <= 10 * tol


@pytest.mark.parametrize("storage", ["f16dev", "bf16dev"])
def test_two_step_16bit_reproducible_and_mode_independent(storage):
    """16-bit two-step runs repeat bit for bit, whichever way the tiles
    holding solids are swept."""
    from pylabolt_amd.solvers.fluidLB import Solver
    pops = []
    for mode in ("sweep", "cells", "sweep"):
        s = Solver(simulation=CASES["cylinder"](), storage=storage,
                   graph_steps=8, verbose=False, two_step=True,
                   solid_tiles=mode)
        s.set_backend()
        assert s.engine.two_step
        s.advance(23)
        pops.append(s.engine.pop())
    assert np.array_equal(pops[0], pops[1])
    assert np.array_equal(pops[0], pops[2])


def test_two_step_force_field_bitwise():
    """A per-cell force field (the adjoint's forcing) through the two-step
    sweep, FP64, Guo forcing, bounce-back walls reached by deep tiles:
    bitwise equal to single steps."""
    from pylabolt_amd.base.collision_operator import CollisionOperator
    from pylabolt_amd.base.state import State
    from pylabolt_amd.engine import D2Q9Engine, elements_from_boundary
    from pylabolt_amd.solvers.fluidLB import FluidLB
    sim, force, _, _ = cases.adjoint_channel(nx=96, ny=256)
    st = State(sim)
    co = CollisionOperator(sim, FluidLB(), st, verbose=False)
    nx, ny = 96, 256
    ff = np.zeros((nx + 2, ny + 2, 2))
    ff[1:-1, 1:-1] = force * 50.0
    lat = st.lattice
    pops = []
    for two in (False, True):
        e = D2Q9Engine(nx, ny, st.fields.solid, st.fields.density,
                       st.fields.velocity, elements_from_boundary(st.boundary),
                       st.boundary.x_periodic, st.boundary.y_periodic, "BGK",
                       co.forcing_fluid, co.omega_fluid, (0.0, 0.0),
                       lat.inv_cs_2, lat.inv_cs_4, lat.weights,
                       st.control.float_min, storage="f64",
                       force_field=ff.reshape(-1, 2), graph_steps=8,
                       two_step=two)
        assert e.two_step == two
        e.advance(31)
        pops.append((e.pop(), e.moments_host()[1]))
    assert np.array_equal(pops[0][0], pops[1][0])
    assert np.array_equal(pops[0][1], pops[1][1])


def test_two_step_default_threshold():
    """By default the two-step sweep runs in FP64 and FP32 storage when the
    two population buffers exceed TWO_STEP_MIN_BYTES: not on the 1024x256
    FP64 channel (54 MiB), on the 2048x512 FP64 (180 MiB) and 2048x1024
    FP32 (180 MiB) ones.  In 16-bit storage it changes the rounding, so it
    runs only when asked, whatever the size (2048x1024 FP16, 90 MiB)."""
    from pylabolt_amd.engine import TWO_STEP_MIN_BYTES
    from pylabolt_amd.solvers.fluidLB import Solver
    for nx, ny, storage, big, want in (
            (1024, 256, "f64", False, False),
            (2048, 512, "f64", True, True),
            (2048, 1024, "f32", True, True),
            (2048, 1024, "f16dev", True, False)):
        s = Solver(simulation=cases.poiseuille_channel(nx, ny),
                   storage=storage, verbose=False)
        s.set_backend()
        e = s.engine
        nbytes = 2 * e.f[0].numel() * e.f[0].element_size()
        assert (nbytes > TWO_STEP_MIN_BYTES) == big
        assert e.two_step == want, (nx, ny, storage, nbytes)
        s.advance(6)
