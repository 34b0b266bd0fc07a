"""Host logic (CPU only): case parsing with the reference's messages, index
conventions, initial fields, obstacle rasterisation, and the C ABI library
(loads, exports every symbol include/d2q9.h declares, struct layout)."""
import ctypes
import os
import re
from types import SimpleNamespace

import numpy as np
import pytest

from pylabolt_amd import cases
from pylabolt_amd.base.boundary import Boundary
from pylabolt_amd.base.domain import Domain
from pylabolt_amd.base.state import State
from pylabolt_amd.utils.helpers import SingleRankComm

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


class Comm:
    def __init__(self, rank, size):
        self.rank, self.size = rank, size

    def Get_rank(self):
        return self.rank

    def Get_size(self):
        return self.size


def _sim(**over):
    sim = cases.lid_driven_cavity(10, 8)
    for k, v in over.items():
        setattr(sim, k, v)
    return sim


# ---- configuration errors (reference messages) ---------------------------

@pytest.mark.parametrize("mutate,msg", [
    (lambda s: delattr(s, "control_dict"), "control_dict not found"),
    (lambda s: s.control_dict.pop("end_time"), "end_time missing in control_dict"),
    (lambda s: s.control_dict.update(precision="quad"), "unsupported precision"),
    (lambda s: s.mesh_dict.update(grid=[0, 4]), "grid dimensions cannot be zero"),
    (lambda s: s.mesh_dict.update(grid=(4, 4)), "must be a list"),
    (lambda s: s.lattice_dict.update(lattice_type="D3Q19"), "Unsupported lattice type"),
    (lambda s: s.transport_dict.update(kin_visc="x"), "kin_visc must be a float/int"),
    (lambda s: s.decompose_dict.update(nx=2), "invalid domain decomposition"),
    (lambda s: s.boundary_dict["walls"].pop("wall"), "wall missing in boundary: walls"),
    (lambda s: s.boundary_dict["walls"].update(segments=[]), "segments cannot be an empty list"),
    (lambda s: s.boundary_dict["walls"].update(segments=[[[3, 0], [1, 0]]]),
     "x2 >= x1 and y2 >= y1"),
    (lambda s: s.boundary_dict["walls"].update(segments=[[[0, 0], [3, 3]]]),
     "axis-aligned"),
    (lambda s: s.boundary_dict["walls"]["fluid"].update(type="slip"),
     "Unsupported boundary condition for fluid: slip"),
    (lambda s: s.boundary_dict["lid"]["fluid"].update(value=0.1),
     "value for fixed_velocity must be a list"),
    (lambda s: s.initial_fields_dict["default"]["fluid"].pop("density"),
     "is missing in default"),
])
def test_config_errors_carry_reference_messages(mutate, msg):
    sim = _sim()
    mutate(sim)
    with pytest.raises(ValueError, match=msg):
        State(sim)


def test_collision_dict_validation():
    from pylabolt_amd.base.collision_operator import CollisionOperator
    from pylabolt_amd.solvers.fluidLB import FluidLB
    sim = _sim()
    st = State(sim)
    sim.collision_dict["fluid"]["model"] = "TRT"
    with pytest.raises(ValueError, match="Unsupported fluid collision model: TRT"):
        CollisionOperator(sim, FluidLB(), st)
    sim.collision_dict["fluid"]["model"] = "BGK"
    sim.collision_dict["fluid"]["forcing_model"] = "shan_chen"
    with pytest.raises(ValueError, match="Unsupported fluid forcing model"):
        CollisionOperator(sim, FluidLB(), st)


def test_omega_matches_reference_formula():
    from pylabolt_amd.base.collision_operator import CollisionOperator
    from pylabolt_amd.solvers.fluidLB import FluidLB
    sim = cases.lid_driven_cavity(128, 128)
    st = State(sim)
    co = CollisionOperator(sim, FluidLB(), st)
    assert st.transport.kin_visc == np.float64(0.127)
    assert co.tau_fluid == 0.8809999999999999
    assert co.omega_fluid == 1.135073779795687


# ---- domain and index conventions ------------------------------------------

@pytest.mark.parametrize("gx,gy,nx,ny", [(10, 7, 1, 1), (10, 7, 3, 2), (9, 9, 3, 3)])
def test_domain_decomposition_covers_every_node_once(gx, gy, nx, ny):
    """Reference tests/unit/test_domain.py:125-216."""
    owned = np.zeros((gx, gy), dtype=int)
    for rank in range(nx * ny):
        sim = SimpleNamespace(decompose_dict={"nx": nx, "ny": ny})
        mesh = SimpleNamespace(grid_global_shape=np.array([gx, gy]))
        d = Domain(sim, mesh, Comm(rank, nx * ny))
        assert tuple(d.shape) == (d.Nx_rank + 2, d.Ny_rank + 2)
        ox, oy = d.offset
        owned[ox:ox + d.Nx_rank, oy:oy + d.Ny_rank] += 1
    assert np.all(owned == 1)


def test_boundary_node_indices_reference_fixture():
    """Reference tests/unit/test_boundary.py:567-588: node of global (i, j)
    is (i + 1) * (Ny + 2) + (j + 1)."""
    Nx, Ny = 7, 5
    sim = cases.lid_driven_cavity(Nx, Ny)
    segs = [[[0, Ny - 1], [Nx - 1, Ny - 1]], [[0, 0], [Nx - 1, 0]],
            [[0, 0], [0, Ny - 1]], [[Nx - 1, 0], [Nx - 1, Ny - 1]]]
    sim.boundary_dict = {"options": {}, "check": {
        "wall": True, "segments": segs, "fluid": {"type": "bounce_back"}}}
    st = State(sim)
    for n, seg in enumerate(segs):
        want = [(i + 1) * (Ny + 2) + (j + 1)
                for i in range(seg[0][0], seg[1][0] + 1)
                for j in range(seg[0][1], seg[1][1] + 1)]
        assert list(st.boundary.boundary_elements[n].boundary_nodes) == want
    locs = [e.location for e in st.boundary.boundary_elements]
    assert locs == ["top", "bottom", "left", "right"]
    e = st.boundary.boundary_elements[1]
    assert list(e.out_list) == [4, 7, 8] and list(e.inv_list) == [2, 5, 6]


def test_periodic_pair_validation():
    sim = cases.poiseuille_channel(8, 6)
    st = State(sim)
    assert st.boundary.x_periodic and not st.boundary.y_periodic
    sim.boundary_dict["inlet_outlet"]["segments"] = [[[0, 0], [0, 4]],
                                                     [[7, 0], [7, 5]]]
    with pytest.raises(ValueError, match="must span the entire y-direction"):
        State(sim)


def test_init_fields_fixed_on_non_ghost_only():
    """Reference tests/unit/test_init_fields.py:226-274."""
    sim = _sim()
    sim.initial_fields_dict = {"default": {"fluid": {
        "velocity": {"type": "fixed", "value": [1, 2]},
        "density": {"type": "fixed", "value": 3},
        "pressure": {"type": "fixed", "value": 4}}}}
    f = State(sim).fields
    assert np.all(f.velocity[~f.ghost_node] == [1, 2])
    assert np.all(f.velocity[f.ghost_node] == [0, 0])
    assert np.all(f.density[~f.ghost_node] == 3)
    assert np.all(f.density[f.ghost_node] == 0)
    assert np.all(f.pressure[~f.ghost_node] == 4)


def test_init_fields_func_uses_global_indices():
    sim = _sim()
    sim.initial_fields_dict["default"]["fluid"]["density"] = {
        "type": "func", "func": lambda i, j: 1.0 + i + 100 * j}
    st = State(sim)
    nxp, nyp = st.domain.shape
    rho = st.fields.density.reshape(nxp, nyp)
    assert rho[1, 1] == 1.0 and rho[3, 2] == 1.0 + 2 + 100
    assert rho[0, 0] == 0.0


@pytest.mark.parametrize("center", [[4.0, 3.0], [0.5, 5.2]])
def test_circle_rasterisation_reference_fixture(center):
    """Reference tests/unit/test_obstacle.py:329-370 (minimum image across
    periodic sides, ghost nodes never solid)."""
    sim = cases.poiseuille_channel(12, 9)
    sim.boundary_dict["walls"]["fluid"]["type"] = "periodic"
    sim.obstacle_dict = {"options": {}, "c": {
        "type": "circle", "radius": 2.5, "center": center, "density": 1.7,
        "static": True}}
    st = State(sim)
    Nx, Ny = 12, 9
    Nxp, Nyp = Nx + 2, Ny + 2
    solid = np.zeros(Nxp * Nyp, dtype=bool)
    for ind in range(Nxp * Nyp):
        i, j = divmod(ind, Nyp)
        rx = i - (center[0] + 1)
        ry = j - (center[1] + 1)
        rxm, rym = rx, ry
        if abs(rx + Nx) < abs(rxm):
            rxm = rx + Nx
        if abs(rx - Nx) < abs(rxm):
            rxm = rx - Nx
        if abs(ry + Ny) < abs(rym):
            rym = ry + Ny
        if abs(ry - Ny) < abs(rym):
            rym = ry - Ny
        if np.sqrt(rxm * rxm + rym * rym) <= 2.5 and not st.fields.ghost_node[ind]:
            solid[ind] = True
    assert np.array_equal(st.fields.solid, solid)
    assert np.all(st.fields.density[solid] == 1.7)
    assert np.all(st.fields.solid_id[solid] == 0)
    assert np.all(st.fields.solid_id[~solid] == -1)


def _two_circles_periodic():
    sim = cases.poiseuille_channel(30, 22, model="BGK", forcing=None)
    sim.obstacle_dict = {
        "options": {},
        "edge": {"type": "circle", "radius": 4.0, "center": [1.0, 10.3],
                 "density": 1.0, "static": True},
        "mid": {"type": "circle", "radius": 3.5, "center": [16.2, 11.0],
                "density": 1.0, "static": True}}
    return sim


@pytest.mark.parametrize("make", [
    lambda: cases.cylinder(40, 24, diameter=8, xc=12),
    _two_circles_periodic])
def test_obstacle_boundary_fields_match_oracle(make):
    """solid_boundary, fluid_boundary, solid_id and the circle normals the
    obstacle operator marks at set-up (saved in every t_<step>.npz) equal
    the oracle's restatement of compute_obstacle_boundary and
    compute_normals_circle, here also across a periodic edge."""
    from oracle import d2q9_oracle as O
    from pylabolt_amd.base.operators import mark_obstacle_boundaries
    st = State(make())
    f = st.fields
    circles = [(o.id, o.center) for o in st.obstacle.obstacles]
    want = O.obstacle_boundary_fields(
        tuple(int(v) for v in st.domain.shape), f.solid, f.ghost_node,
        f.solid_id, circles, st.boundary.x_periodic, st.boundary.y_periodic)
    mark_obstacle_boundaries(st)
    assert f.solid_boundary.any() and f.fluid_boundary.any()
    assert np.array_equal(f.solid_boundary, want[0])
    assert np.array_equal(f.fluid_boundary, want[1])
    assert np.array_equal(f.solid_id, want[2])
    assert np.array_equal(f.surface_normals, want[3])


def test_single_rank_comm_abort_raises():
    with pytest.raises(SystemExit):
        SingleRankComm().Abort()


def test_load_simulation_from_directory(tmp_path):
    from pylabolt_amd.utils.helpers import load_simulation
    (tmp_path / "simulation.py").write_text(
        "control_dict = {'start_time': 0}\n")
    mod = load_simulation(path=tmp_path)
    assert mod.control_dict["start_time"] == 0
    with pytest.raises(ImportError, match="Missing simulation.py"):
        load_simulation(path=tmp_path / "nowhere")


# ---- the C ABI library (no GPU calls) -------------------------------------

def _header_functions():
    text = open(os.path.join(ROOT, "include", "d2q9.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(d2q9_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from pylabolt_amd import _native
    lib = _native.load_library()      # also checks the struct layout
    names = _header_functions()
    assert len(names) >= 15
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) == set(_native.SIGNATURES)


def test_abi_layout_matches_ctypes_mirror():
    from pylabolt_amd import _native
    lib = _native.load_library()
    assert _native.abi_mismatches(lib) == []


def test_stale_library_is_refused(tmp_path):
    """The library carries the digest of the sources it was built from; a
    tree whose sources differ (here: one byte of one kernel source) refuses
    it, the unchanged tree loads it."""
    import shutil
    from pylabolt_amd import _native, build
    csrc = tmp_path / "csrc"
    shutil.copytree(build.CSRC, csrc,
                    ignore=shutil.ignore_patterns("*.o", "*.so"))
    header = tmp_path / "d2q9.h"
    shutil.copy(build.HEADER, header)
    _native.load_library(csrc=str(csrc), header=str(header))
    src = csrc / "d2q9_step.hip"
    src.write_text(src.read_text() + "\n// edited\n")
    with pytest.raises(_native.NativeError, match="stale"):
        _native.load_library(csrc=str(csrc), header=str(header))
    src.write_text(src.read_text()[:-len("\n// edited\n")])
    header.write_text(header.read_text().replace("cpl = 1", "cpl = 1 "))
    with pytest.raises(_native.NativeError, match="stale"):
        _native.load_library(csrc=str(csrc), header=str(header))


def test_engine_refuses_without_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    from pylabolt_amd import _native
    _native._lib = None
    with pytest.raises(_native.NativeError, match="no HIP device"):
        _native.lib()


def test_force_field_without_guo_forcing_is_refused():
    """A body force acts through the Guo source term: the engine and the
    adjoint refuse a force field with forcing None before touching the GPU
    (the C ABI returns D2Q9_E_UNSUPPORTED for such a descriptor)."""
    from pylabolt_amd import cases
    from pylabolt_amd.adjoint import FlowRollout
    from pylabolt_amd.engine import D2Q9Engine
    w = [4 / 9] + [1 / 9] * 4 + [1 / 36] * 4
    with pytest.raises(ValueError, match="Guo forcing"):
        D2Q9Engine(4, 4, np.zeros(36, bool), np.ones(36), np.zeros((36, 2)),
                   [], True, True, "BGK", None, 1.0, (0.0, 0.0), 3.0, 9.0, w,
                   1e-16, force_field=np.zeros((36, 2)))
    sim, _, _, _ = cases.adjoint_channel(8, 6, steps=4, forcing=None)
    with pytest.raises(ValueError, match="Guo forcing"):
        FlowRollout(sim, 4)


def test_bench_child_leg_returns_record_or_error():
    """The benchmark runs its adjoint leg in a child process: the child's
    record comes back, a child that aborts gives an error record (and the
    parent carries on), and an unused child exits when the pipe closes."""
    import bench
    assert bench.ChildLeg("tests.bench_child_probe:ok").run(timeout=120) == \
        {"ok": 1}
    rec = bench.ChildLeg("tests.bench_child_probe:abort").run(timeout=120)
    assert "error" in rec and "status -6" in rec["error"], rec
    leg = bench.ChildLeg("tests.bench_child_probe:ok")
    leg.close()
    assert leg.proc.returncode == 0
