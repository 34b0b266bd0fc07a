"""FP64 parity of the HIP engine with the oracle (bar: max relative error
<= 1e-12 on pop_fluid, pop_fluid_new, density and velocity; the FP64 path
is expected to be bit-exact and the tests report whether it is).

Every case runs the same steps in the engine (through the C ABI) and in the
NumPy oracle, with the reference's semantics for each boundary type, both
collision models, every forcing, periodic wraps, obstacles, plain and
flagged tiles, and HIP-graph replay.
"""
import numpy as np
import pytest

from oracle import d2q9_oracle as O
from pylabolt_amd import cases
from tests.helpers import max_rel, oracle_case

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _engine(sim, graph_steps=16, storage="f64"):
    from pylabolt_amd.solvers.fluidLB import Solver
    s = Solver(simulation=sim, storage=storage, graph_steps=graph_steps,
               verbose=False)
    s.set_backend()
    return s


def _compare(solver, case, label):
    eng = solver.engine
    errs = {}
    pop = eng.pop()
    pop_new = eng.pop_new()
    rho, vel = eng.moments_host()
    fl = case.fluid
    errs["pop"] = max_rel(pop[fl], case.pop[fl])
    errs["pop_new"] = max_rel(pop_new[fl], case.pop_new[fl])
    errs["density"] = max_rel(rho[fl], case.density[fl])
    errs["velocity"] = max_rel(vel[fl], case.velocity[fl])
    exact = all(v == 0.0 for v in errs.values())
    print(f"[parity] {label}: {errs} bitwise={exact}", flush=True)
    for k, v in errs.items():
        assert v <= TOL, f"{label}: {k} max rel err {v}"
    return exact


SMALL = {
    "cavity_bgk": lambda: cases.lid_driven_cavity(12, 9),
    "cavity_mrt": lambda: cases.lid_driven_cavity(11, 10, model="MRT"),
    "channel_bgk_guo2": lambda: cases.poiseuille_channel(
        10, 7, model="BGK", forcing="guo_second_order"),
    "channel_bgk_guo1": lambda: cases.poiseuille_channel(
        9, 8, model="BGK", forcing="guo_linear"),
    "channel_mrt_guo2": lambda: cases.poiseuille_channel(10, 7),
    "channel_mrt_guo1": lambda: cases.poiseuille_channel(
        8, 6, forcing="guo_linear"),
    "cylinder_zg_pressure": lambda: cases.cylinder(40, 20, diameter=8, xc=12),
    # wide enough (ny >= 64) for bulk tiles whose boundary, near-solid and
    # stored-moment cells go through the fix-up pass
    "cavity_wide": lambda: cases.lid_driven_cavity(9, 192),
    "channel_wide_mrt": lambda: cases.poiseuille_channel(12, 128),
    "cylinder_wide": lambda: cases.cylinder(44, 130, diameter=10, xc=14),
}


@pytest.mark.parametrize("name", sorted(SMALL))
@pytest.mark.parametrize("graph_steps", [2, 4])
def test_small_cases_exact(name, graph_steps):
    sim = SMALL[name]()
    if name.startswith("channel"):
        sim.forcing_dict["gravity"] = [2e-5, 3e-6]
    case = oracle_case(sim)
    O.initialize_pop(case)
    solver = _engine(sim, graph_steps=graph_steps)
    done = 0
    for chunk in (1, 1, 5, 13):
        solver.advance(chunk)
        O.run(case, chunk)
        done += chunk
        _compare(solver, case, f"{name} graph={graph_steps} step={done}")


def test_single_time_step_api():
    sim = cases.lid_driven_cavity(16, 12)
    case = oracle_case(sim)
    O.initialize_pop(case)
    solver = _engine(sim)
    for _ in range(3):
        solver.single_time_step()
    O.run(case, 3)
    solver.sync_fields()
    f = solver.state.fields
    fl = case.fluid
    assert max_rel(f.pop_fluid[fl], case.pop[fl]) <= TOL
    assert max_rel(f.pop_fluid_new[fl], case.pop_new[fl]) <= TOL
    assert max_rel(f.density[fl], case.density[fl]) <= TOL
    assert max_rel(f.velocity[fl], case.velocity[fl]) <= TOL


def test_unfused_operator_verbs_match():
    """compute_fields -> collide -> halo -> stream -> set_boundary through
    the operator objects equals the fused step."""
    sim = cases.poiseuille_channel(12, 9, model="BGK")
    sim.forcing_dict["gravity"] = [1e-5, 0.0]
    case = oracle_case(sim)
    O.initialize_pop(case)
    s = _engine(sim)
    st, be = s.state, s.backend
    for _ in range(4):
        s.compute_fields_operator.compute_fields(st, be, field=["density"])
        s.force_operator.compute_force_field(st, be)
        s.collision_operator.collide(st, be, fluid=True)
        s.mpi_operator.halo_exchange(st, be, float_buffers=["pop_fluid"])
        s.streaming_operator.stream(st, be, fluid=True)
        s.boundary_operator.set_boundary(st, be, fluid=True)
    O.run(case, 4)
    _compare(s, case, "unfused verbs")


def test_graph_replay_long_periodic_plain_tiles():
    """A doubly periodic forced box large enough for plain tiles, many HIP
    graph replays between checks."""
    sim = cases.poiseuille_channel(70, 300, model="BGK",
                                   forcing="guo_second_order")
    sim.boundary_dict = {
        "options": {},
        "x": {"wall": False, "segments": [[[0, 0], [0, 299]],
                                          [[69, 0], [69, 299]]],
              "fluid": {"type": "periodic"}},
        "y": {"wall": False, "segments": [[[0, 0], [69, 0]],
                                          [[0, 299], [69, 299]]],
              "fluid": {"type": "periodic"}},
    }
    sim.forcing_dict["gravity"] = [1e-5, -2e-6]
    sim.initial_fields_dict["default"]["fluid"]["velocity"] = {
        "type": "func", "func": lambda i, j: (0.01 * np.sin(j / 7.0),
                                              0.02 * np.cos(i / 5.0))}
    case = oracle_case(sim)
    O.initialize_pop(case)
    solver = _engine(sim, graph_steps=8)
    assert (solver.engine.tables.tile_class == 0).any()
    solver.advance(41)
    O.run(case, 41)
    _compare(solver, case, "periodic box 70x300 41 steps")


def test_cavity_128_long_vs_cpp(ref_lib):
    """The cavity workload at its real size against the C++ restatement."""
    sim = cases.lid_driven_cavity(128, 128)
    case = oracle_case(sim)
    r = ref_lib.RefRunner(case)
    r.initialize()
    solver = _engine(sim)
    n = 1000
    solver.advance(n)
    r.run(n)
    _compare(solver, case, f"cavity 128x128 {n} steps")


def test_mass_fp64_matches_host_sum():
    sim = cases.lid_driven_cavity(40, 30)
    solver = _engine(sim)
    solver.advance(25)
    m = solver.engine.mass()
    pop = solver.engine.pop()
    assert abs(m - pop.sum()) <= 1e-11 * abs(m)


def test_cylinder_full_size_vs_cpp(ref_lib):
    """The cylinder workload at its full 8192 x 2048 size (obstacle,
    fixed-pressure ends, zero-gradient sides) against the C++ restatement
    after a few dozen steps."""
    sim = cases.cylinder()
    case = oracle_case(sim)
    r = ref_lib.RefRunner(case)
    r.initialize()
    solver = _engine(sim, graph_steps=8)
    n = 25
    solver.advance(n)
    r.run(n)
    _compare(solver, case, f"cylinder 8192x2048 {n} steps")


def test_channel_full_size_conserves_mass():
    """Periodic Guo-forced channel at 4096 x 1024 (bounce-back walls): the
    total mass is constant up to FP64 round-off over 1,000 steps (the Guo
    source sums to zero over the directions)."""
    sim = cases.poiseuille_channel(4096, 1024, model="BGK",
                                   forcing="guo_second_order")
    solver = _engine(sim)
    eng = solver.engine
    m0 = eng.mass()
    solver.advance(1000)
    m1 = eng.mass()
    rel = abs(m1 - m0) / m0
    print(f"[mass] channel 4096x1024 1000 steps: {m0!r} -> {m1!r} rel {rel:.2e}",
          flush=True)
    assert rel <= 1e-12



def test_mrt_rates_outside_the_reference_form_are_refused():
    """The kernels evaluate the reference's S = (1,..,1,omega,omega): the
    C ABI returns D2Q9_E_UNSUPPORTED (-3) for a descriptor with another
    rate on rows 0..6, before any launch, takes other stress rates (rows 7
    and 8), and the engine refuses such rates with ValueError."""
    import ctypes
    from pylabolt_amd import _native as nt
    from pylabolt_amd.solvers.fluidLB import Solver
    sim = cases.lid_driven_cavity(12, 10, model="MRT")
    eng = _engine(sim).engine

    def step(L):
        return eng.lib.d2q9_fused_step(
            ctypes.byref(L), nt.ptr(eng.f[0]), nt.ptr(eng.f[1]),
            nt.ptr(eng.mom[0]), nt.ptr(eng.mom[1]), 0, eng._stream_ptr)
    bad = nt.Lattice.from_buffer_copy(eng.L)
    bad.rates[5] = bad.omega
    assert step(bad) == -3
    assert b"rows 0..6" in eng.lib.d2q9_last_error()
    ok = nt.Lattice.from_buffer_copy(eng.L)
    ok.rates[7] = ok.rates[8] = 1.5
    ok.cf.rates[7] = ok.cf.rates[8] = 1.5
    assert step(ok) == 0
    eng.synchronize()
    s = Solver(simulation=sim, verbose=False)
    s.collision_operator.S[5] = s.collision_operator.omega_fluid
    with pytest.raises(ValueError, match="rows 0..6"):
        s.set_backend()


def test_mrt_free_stress_rates_bitwise(monkeypatch):
    """Stress rates s7 = 1.3, s8 = 1.7 (the C ABI leaves rows 7 and 8
    free): the engine against the oracle's reduced MRT with those rates,
    forced channel with a y force, 30 steps, bitwise bar 1e-12."""
    from pylabolt_amd.solvers.fluidLB import Solver
    sim = cases.poiseuille_channel(24, 16, model="MRT",
                                   forcing="guo_second_order")
    sim.forcing_dict["gravity"] = [3e-5, 1e-5]
    s = Solver(simulation=sim, verbose=False)
    s.collision_operator.S[7], s.collision_operator.S[8] = 1.3, 1.7
    s.set_backend()
    case = oracle_case(sim)
    good = O.mrt_collide
    monkeypatch.setattr(O, "mrt_collide",
                        lambda c: good(c, stress_rates=(1.3, 1.7)))
    O.initialize_pop(case)
    s.advance(30)
    O.run(case, 30)
    _compare(s, case, "MRT s7=1.3 s8=1.7")
