"""Solver.run end to end: residues, status lines, field files, metadata.

The reference run loop (solvers/fluidLB.py:309-395) computes residues and
prints a status line every ``std_out_interval`` steps and writes
``output/fields/t_<step>.npz`` every ``save_interval`` steps, after a
``metadata.json`` that its reconstruct/npz2vtk tools read.  Checked here on
a cylinder (solid, boundary-node and normal fields are non-trivial):

* residues against the formula of compute_residue_scalar/vector
  (parallel/cpu/compute_residues_kernels.py:6-73, utils/residues.py:171-222)
  applied on the host to the oracle's moments, with phi_old starting at
  zero (residues.py:53-64), to 1e-12 relative (the GPU sums in another
  order);
* status lines in the format of utils/helpers.py:29-56;
* field files: the keys of FluidLB.save_fields, inner shapes and dtypes,
  inner index (i - 1) * Ny + (j - 1), density/velocity equal to the
  oracle's (bitwise: the FP64 engine is), the obstacle boundary fields equal
  to the oracle's restatement of compute_obstacle_boundary;
* metadata.json with the keys of utils/io_operator.py:96-140.
"""
import json

import numpy as np
import pytest

from oracle import d2q9_oracle as O
from pylabolt_amd import cases
from pylabolt_amd.base.state import State
from tests.helpers import oracle_case

pytestmark = pytest.mark.gpu

NX, NY, END, STD_OUT, SAVE = 40, 24, 20, 5, 10


def _sim():
    sim = cases.cylinder(NX, NY, diameter=8, xc=12)
    sim.control_dict.update(end_time=END, std_out_interval=STD_OUT,
                            save_interval=SAVE)
    return sim


def _inner(a, nxp, nyp):
    if a.ndim == 1:
        return a.reshape(nxp, nyp)[1:-1, 1:-1].reshape(-1)
    return a.reshape(nxp, nyp, a.shape[1])[1:-1, 1:-1].reshape(-1, a.shape[1])


def _oracle_run():
    """Oracle moments at every step; residues at the std_out steps and the
    fields at the save steps."""
    case = oracle_case(_sim())
    O.initialize_pop(case)
    fl = case.fluid
    fm = np.finfo(np.float64).eps
    old_r = np.zeros(case.size)
    old_u = np.zeros((case.size, 2))
    res, saved = {}, {0: (case.density.copy(), case.velocity.copy())}
    for t in range(1, END + 1):
        O.single_time_step(case)
        if t % STD_OUT == 0:
            r, u = case.density[fl], case.velocity[fl]
            dr, du = r - old_r[fl], u - old_u[fl]
            res[t] = (np.sqrt((dr * dr).sum() / ((old_r[fl] ** 2).sum() + fm)),
                      np.sqrt((du[:, 0] ** 2).sum() /
                              ((old_u[fl, 0] ** 2).sum() + fm)),
                      np.sqrt((du[:, 1] ** 2).sum() /
                              ((old_u[fl, 1] ** 2).sum() + fm)))
            old_r[fl], old_u[fl] = r, u
        if t % SAVE == 0:
            saved[t] = (case.density.copy(), case.velocity.copy())
    return case, res, saved


def test_run_residues_status_lines_and_field_output(tmp_path):
    from pylabolt_amd.solvers.fluidLB import FluidLB, Solver
    s = Solver(simulation=_sim(), verbose=False, output_root=str(tmp_path))
    s.set_backend()
    seen = []
    log = s.logger.log_data

    def record(state, t, **values):
        seen.append((t, np.concatenate([np.array(v, dtype=np.float64)
                                        for v in values.values()])))
        log(state, t, **values)
    s.logger.log_data = record
    s.run()
    case, res, saved = _oracle_run()

    # residues and status lines
    assert [t for t, _ in seen] == list(range(STD_OUT, END + 1, STD_OUT))
    for t, got in seen:
        want = np.array(res[t])
        rel = np.abs(got - want) / np.abs(want)
        print(f"[run] t={t} residues {got} oracle {want} rel {rel.max():.1e}",
              flush=True)
        assert rel.max() <= 1e-12, (t, got, want)
    assert len(s.logger.lines) == END // STD_OUT
    for line, (t, got) in zip(s.logger.lines, seen):
        parts = [p.strip() for p in line.split("|")]
        assert parts[0] == f"{'time:':<5} {t:<10}".strip()
        assert parts[1] == f"{'res_density':<5}: {got[0]:.5e}"
        assert parts[2] == (f"{'res_velocity':<5}: "
                            f"({got[1]:.5e}, {got[2]:.5e})")

    # field files
    fdir = tmp_path / "output" / "fields"
    assert sorted(p.name for p in fdir.iterdir()) == \
        ["t_0.npz", "t_10.npz", "t_20.npz"]
    nxp, nyp = NX + 2, NY + 2
    st = State(_sim())
    f = st.fields
    circles = [(o.id, o.center) for o in st.obstacle.obstacles]
    sb, fb, sid, nrm = O.obstacle_boundary_fields(
        (nxp, nyp), f.solid, f.ghost_node, f.solid_id, circles,
        st.boundary.x_periodic, st.boundary.y_periodic)
    want_static = {"solid": f.solid, "solid_id": sid, "solid_boundary": sb,
                   "fluid_boundary": fb, "surface_normals": nrm}
    names = FluidLB().save_fields
    for t, (rho, vel) in saved.items():
        with np.load(fdir / f"t_{t}.npz") as z:
            assert sorted(z.files) == sorted(names)
            assert z["density"].shape == (NX * NY,)
            assert z["velocity"].shape == (NX * NY, 2)
            assert z["density"].dtype == np.float64
            assert z["solid"].dtype == np.bool_
            assert z["solid_id"].dtype == np.int64
            fl = _inner(case.fluid, nxp, nyp)
            assert np.array_equal(z["density"][fl], _inner(rho, nxp, nyp)[fl])
            assert np.array_equal(z["velocity"][fl], _inner(vel, nxp, nyp)[fl])
            for n, w in want_static.items():
                assert np.array_equal(z[n], _inner(np.asarray(w), nxp, nyp)), n
    # inner index (i - 1) * Ny + (j - 1): the solid cell nearest the centre
    with np.load(fdir / "t_10.npz") as z:
        o = st.obstacle.obstacles[0]
        i, j = int(round(o.center[0])), int(round(o.center[1]))
        assert z["solid"][i * NY + j] and z["solid_id"][i * NY + j] == o.id
        assert not z["solid"][0]

    # metadata.json (utils/io_operator.py:96-140)
    meta = json.loads((tmp_path / "metadata.json").read_text())
    assert set(meta) == {"pylabolt", "control", "mesh", "decomposition",
                         "fields_saved", "engine"}
    assert meta["engine"] == {"storage": "f64", "two_step": False}
    assert meta["pylabolt"]["solver"] == "fluidLB"
    assert meta["control"] == {"end_time": END, "start_time": 0,
                               "save_interval": SAVE,
                               "checkpoint_interval": None}
    assert meta["mesh"] == {"size": NX * NY, "shape": [NX, NY]}
    assert meta["decomposition"] == {"nx": 1, "ny": 1}
    assert list(meta["fields_saved"]) == names
    assert meta["fields_saved"]["velocity"] == {"components": 2,
                                                "dtype": "float64"}
    assert meta["fields_saved"]["solid"] == {"components": 1, "dtype": "bool"}
