"""Momentum-exchange forces on obstacles and walls.

CPU: the oracle's restatement of force_torque_kernels.py is pinned by the
momentum balance of one streaming pass -- with halfway bounce-back walls,
periodic sides and a static obstacle, the fluid's total momentum before
streaming minus after streaming equals the summed link forces -- and a
planted sign error fails it; the host's node selection equals the oracle's
restatement of the reference's boundary-node marking.
GPU: the engine's FP64 forces and torque equal the oracle's on the same
state (bar 1e-12 relative + 1e-13 absolute: the link values are O(0.1) and
the sums run in another order), satisfy the same balance, and the solver
writes the reference's history files.
"""
import os

import numpy as np
import pytest

from oracle import d2q9_oracle as O
from pylabolt_amd import cases
from pylabolt_amd.base.operators import obstacle_boundary_nodes
from pylabolt_amd.base.state import State
from tests.helpers import oracle_case


def channel_with_cylinder(nx=48, ny=30, interval=1):
    sim = cases.poiseuille_channel(nx, ny, model="BGK",
                                   forcing="guo_second_order")
    sim.forcing_dict["gravity"] = [4e-5, 1e-6]
    sim.boundary_dict["options"] = {"compute_force": True,
                                    "write_boundary_data":
                                        {"interval": interval}}
    sim.obstacle_dict = {
        "options": {"compute_force_torque": True,
                    "write_obstacle_data": {"interval": interval}},
        "cyl": {"type": "circle", "radius": 4.0, "center": [15.0, 14.3],
                "density": 1.0, "static": True}}
    return sim


def _momentum(pop, fluid):
    cx = O.CX.astype(np.float64)
    cy = O.CY.astype(np.float64)
    p = pop[fluid]
    return np.array([(p * cx).sum(), (p * cy).sum()])


def _oracle_forces(sim, case):
    st = State(sim)
    fb, sid = O.fluid_boundary_nodes(case.shape, case.solid, case.ghost,
                                     st.fields.solid_id)
    walls = [O.boundary_force(case, e) for e, be in
             zip(case.elements, st.boundary.boundary_elements) if be.wall]
    obst = [O.obstacle_force_torque(case, sid, fb, o.center, o.id)
            for o in st.obstacle.obstacles]
    return walls, obst


@pytest.mark.parametrize("steps", [0, 7])
def test_oracle_forces_balance_streaming_momentum(steps):
    sim = channel_with_cylinder()
    case = oracle_case(sim)
    O.initialize_pop(case)
    O.run(case, steps)
    walls, obst = _oracle_forces(sim, case)
    total = sum(w for w in walls) + sum(o[:2] for o in obst)
    dp = _momentum(case.pop, case.fluid) - _momentum(case.pop_new, case.fluid)
    assert np.abs(total - dp).max() <= 1e-13 * max(1.0, np.abs(dp).max())
    assert np.abs(total).max() > 0 or steps == 0


def test_planted_force_sign_is_caught(monkeypatch):
    """Adding instead of subtracting the returning population breaks the
    balance."""
    sim = channel_with_cylinder()
    case = oracle_case(sim)
    O.initialize_pop(case)
    O.run(case, 5)
    wrong = case.pop_new.copy()
    # a kernel that summed pop[k] c_k + pop_new[inv k] c_{inv k} equals the
    # restatement fed with -pop_new
    walls = [O.boundary_force(case, e, pop_new=-wrong) for e in case.elements
             if e["type"] == "bounce_back"]
    total = sum(walls)
    dp = _momentum(case.pop, case.fluid) - _momentum(case.pop_new, case.fluid)
    assert np.abs(total - dp).max() > 1e-6


def test_host_boundary_nodes_match_oracle():
    sim = channel_with_cylinder()
    st = State(sim)
    case = oracle_case(sim)
    fb, sid = O.fluid_boundary_nodes(case.shape, case.solid, case.ghost,
                                     st.fields.solid_id)
    nodes, masks, ids = obstacle_boundary_nodes(st)
    assert sorted(nodes.tolist()) == sorted(np.nonzero(fb)[0].tolist())
    assert np.array_equal(ids, sid[nodes])
    nyp = case.shape[1]
    for n, m in zip(nodes, masks):
        x, y = divmod(int(n), nyp)
        want = sum(1 << k for k in range(9)
                   if case.solid[(x + O.CX[k]) * nyp + y + O.CY[k]])
        assert int(m) == want


@pytest.mark.gpu
@pytest.mark.parametrize("steps", [0, 1, 9])
def test_engine_forces_match_oracle(steps):
    from pylabolt_amd.solvers.fluidLB import Solver
    sim = channel_with_cylinder()
    case = oracle_case(sim)
    O.initialize_pop(case)
    O.run(case, steps)
    walls, obst = _oracle_forces(sim, case)
    s = Solver(simulation=sim, verbose=False)
    s.set_backend(verbose=False)
    s.advance(steps)
    s.boundary_operator.compute_force(s.state, s.backend, s.mpi_operator,
                                      time_step=steps)
    s.obstacle_operator.compute_force_torque(s.state, s.backend,
                                             s.mpi_operator)
    got_w = [e.force.copy() for e in s.state.boundary.boundary_elements
             if e.wall]
    o = s.state.obstacle.obstacles[0]
    got_o = np.array([o.force[0], o.force[1], o.torque])
    print(f"[forces] step {steps}: walls {got_w} obstacle {got_o}",
          flush=True)
    # the link values are O(0.1); the totals may cancel to 1e-5, so the
    # bar is relative plus an absolute 1e-13 for the order of the sums
    for g, w in zip(got_w, walls):
        assert np.allclose(g, w, rtol=1e-12, atol=1e-13)
    assert np.allclose(got_o, obst[0], rtol=1e-12, atol=1e-13)
    # the same balance on the engine's own arrays
    eng = s.engine
    fl = case.fluid
    dp = _momentum(eng.pop(), fl) - _momentum(eng.pop_new(), fl)
    total = sum(got_w) + got_o[:2]
    assert np.abs(total - dp).max() <= 1e-12 * max(1.0, np.abs(dp).max())


@pytest.mark.gpu
def test_solver_writes_reference_histories(tmp_path):
    from pylabolt_amd.solvers.fluidLB import Solver
    sim = channel_with_cylinder(interval=4)
    sim.control_dict["end_time"] = 12
    sim.control_dict["std_out_interval"] = None
    s = Solver(simulation=sim, verbose=False, output_root=str(tmp_path))
    s.set_backend(verbose=False)
    s.run(verbose=False)
    hist = tmp_path / "output" / "histories"
    obst = (hist / "cyl.dat").read_text().splitlines()
    assert obst[0].startswith("#     PyLaBolt obstacle history")
    rows = [ln.split() for ln in obst if not ln.startswith("#")]
    assert [int(r[0]) for r in rows] == [0, 4, 8, 12]
    assert len(rows[-1]) == 10
    last = s.state.obstacle.obstacles[0]
    assert float(rows[-1][7]) == float(f"{last.force[0]:.16e}")
    walls = sorted(p.name for p in hist.iterdir() if p.name != "cyl.dat")
    assert walls == ["walls_0.dat", "walls_1.dat"]
    rows = [ln.split() for ln in (hist / "walls_0.dat").read_text()
            .splitlines() if not ln.startswith("#")]
    assert [int(r[0]) for r in rows] == [0, 4, 8, 12]


@pytest.mark.gpu
def test_reduced_precision_forces_close_to_fp64():
    from pylabolt_amd.solvers.fluidLB import Solver
    out = {}
    for storage in ("f64", "f32"):
        sim = channel_with_cylinder()
        s = Solver(simulation=sim, storage=storage, verbose=False)
        s.set_backend(verbose=False)
        s.advance(40)
        s.obstacle_operator.compute_force_torque(s.state, s.backend,
                                                 s.mpi_operator)
        o = s.state.obstacle.obstacles[0]
        out[storage] = np.array([o.force[0], o.force[1], o.torque])
    rel = np.abs(out["f32"] - out["f64"]).max() / np.abs(out["f64"]).max()
    print(f"[forces] f32 vs f64 rel {rel:.2e}", flush=True)
    assert rel <= 1e-4
