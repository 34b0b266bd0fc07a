"""Thin operators of the reference API whose work the fused sweep does.

Reference counterparts: base/compute_fields_operator.py,
base/streaming_operator.py, base/boundary_operator.py,
parallel/MPI_operator.py and base/obstacle_operator.py.  They keep the
reference's constructor and verb signatures so that a caller written
against the reference runs unchanged; each verb maps onto the engine:

* ``compute_fields(field=[...])``: rho/u of ``pop_fluid_new``
  (compute_fields_kernels.py:5-65);
* ``stream`` + ``set_boundary``: one stream-and-boundary sweep
  (streaming_kernels.py:5-47, fluid_boundary_kernels.py:5-152);
* ``halo_exchange``: nothing to do on one GPU, the periodic wrap is folded
  into the streaming index (MPI_kernels.py:34-99);
* ``move_obstacles``: static bodies only;
* ``compute_force`` / ``compute_force_torque``: momentum exchange on the
  wall and obstacle links (force_torque_kernels.py:11-141), one FP64
  reduction over the boundary nodes (d2q9_momentum_exchange).
"""
import numpy as np

from pylabolt_amd.base.lattice import CX, CY


def obstacle_boundary_nodes(state):
    """Fluid boundary nodes of the obstacles, their owning obstacle and the
    direction mask of their solid links.

    Reference: compute_obstacle_boundary (parallel/cpu/obstacle_kernels.py,
    the ``elif not solid[ind]`` branch): a non-ghost fluid node with a solid
    neighbour at (i + c_k) is a fluid boundary node and takes the solid_id
    of the first such neighbour in direction order; the force kernel
    (force_torque_kernels.py:45-90) sums every direction whose neighbour is
    solid.  Returns (nodes, masks, ids), nodes as padded indices."""
    nxp, nyp = (int(v) for v in state.domain.shape)
    f = state.fields
    solid = np.asarray(f.solid, bool).reshape(nxp, nyp)
    ghost = np.asarray(f.ghost_node, bool).reshape(nxp, nyp)
    sid = np.asarray(f.solid_id).reshape(nxp, nyp)
    cand = ~solid & ~ghost
    mask = np.zeros((nxp, nyp), dtype=np.uint32)
    first = np.full((nxp, nyp), -1, dtype=np.int64)
    ii, jj = np.nonzero(cand)
    for k in range(9):
        ni, nj = ii + CX[k], jj + CY[k]
        hit = solid[ni, nj]
        mask[ii[hit], jj[hit]] |= np.uint32(1 << k)
        take = hit & (first[ii, jj] < 0)
        first[ii[take], jj[take]] = sid[ni[take], nj[take]]
    sel = mask[ii, jj] != 0
    nodes = ii[sel] * nyp + jj[sel]
    return nodes, mask[ii[sel], jj[sel]], first[ii[sel], jj[sel]]


def mark_obstacle_boundaries(state):
    """Boundary fields of the obstacles, as the reference's ObstacleOperator
    leaves them at set-up (obstacle_operator.py:36-45): the kernels
    compute_obstacle_boundary and compute_normals_circle
    (parallel/cpu/obstacle_kernels.py:97-143, 307-348), vectorised.

    * ``solid_boundary``: a non-ghost solid node with a non-solid neighbour
      at (i + c_k);
    * ``fluid_boundary``: a non-ghost fluid node with a solid neighbour; it
      takes the ``solid_id`` of the first one in direction order;
    * ``surface_normals``: on both kinds of node of a circle, the unit
      vector from the centre to the node (global node coordinates i - 1,
      j - 1, minimum image across periodic sides).

    Raises RuntimeError when two obstacles share a fluid boundary node
    (check_fluid_boundary_overlap, obstacle_operator.py:156-185)."""
    nxp, nyp = (int(v) for v in state.domain.shape)
    f = state.fields
    solid = np.asarray(f.solid, bool).reshape(nxp, nyp)
    ghost = np.asarray(f.ghost_node, bool).reshape(nxp, nyp)
    sid = f.solid_id.reshape(nxp, nyp)
    sb = f.solid_boundary.reshape(nxp, nyp)
    fb = f.fluid_boundary.reshape(nxp, nyp)
    ii, jj = np.nonzero(~ghost)
    sol = solid[ii, jj]
    si, sj = ii[sol], jj[sol]
    fi, fj = ii[~sol], jj[~sol]
    done = np.zeros(fi.size, dtype=bool)
    take_id = np.full(fi.size, -1, dtype=sid.dtype)
    for k in range(9):
        sb[si, sj] |= ~solid[si + CX[k], sj + CY[k]]
        hit = solid[fi + CX[k], fj + CY[k]] & ~done
        take_id[hit] = sid[fi[hit] + CX[k], fj[hit] + CY[k]]
        done |= hit
    fb[fi[done], fj[done]] = True
    sid[fi[done], fj[done]] = take_id[done]
    # overlap: a fluid boundary node with a neighbour of another obstacle
    bi, bj = np.nonzero(fb & ~ghost)
    for k in range(9):
        nb = sid[bi + CX[k], bj + CY[k]]
        if np.any((nb != sid[bi, bj]) & (nb != -1)):
            raise RuntimeError(
                "Fluid Boundary node overlap detected: two solid obstacles "
                "have a common fluid boundary node")
    nrm = f.surface_normals.reshape(nxp, nyp, 2)
    Nx, Ny = (int(v) for v in state.mesh.grid_global_shape)
    off = state.domain.offset
    for o in state.obstacle.obstacles:
        if o.type != "circle":
            continue
        oi, oj = np.nonzero((sid == o.id) & (sb | fb))
        rx = (oi - 1 + off[0]) - o.center[0]
        ry = (oj - 1 + off[1]) - o.center[1]
        if state.boundary.x_periodic:
            rx = np.where(np.abs(rx + Nx) < np.abs(rx), rx + Nx, rx)
            rx = np.where(np.abs(rx - Nx) < np.abs(rx), rx - Nx, rx)
        if state.boundary.y_periodic:
            ry = np.where(np.abs(ry + Ny) < np.abs(ry), ry + Ny, ry)
            ry = np.where(np.abs(ry - Ny) < np.abs(ry), ry - Ny, ry)
        mag = np.sqrt(rx * rx + ry * ry)
        nrm[oi, oj, 0] = rx / mag
        nrm[oi, oj, 1] = ry / mag


def wall_boundary_groups(state):
    """One group per wall element: its non-solid nodes and the out_list
    directions (compute_boundary_force_single_phase,
    force_torque_kernels.py:98-141)."""
    f = state.fields
    solid = np.asarray(f.solid, bool)
    groups = []
    for e in state.boundary.boundary_elements:
        if not e.wall:
            continue
        nodes = np.asarray(e.boundary_nodes, dtype=np.int64)
        nodes = nodes[~solid[nodes]]
        m = 0
        for k in e.out_list:
            m |= 1 << int(k)
        groups.append({"element": e, "nodes": nodes,
                       "masks": np.full(nodes.size, m, dtype=np.uint32),
                       "torque": False})
    return groups


class ComputeFieldsOperator:
    def __init__(self, model, state, verbose=True):
        self.model = model
        self.moment_fields = model.compute_fields_config["moment_fields"]

    def set_backend(self, state, backend, verbose=True):
        return None

    def compile(self, state, backend, verbose=True):
        return None

    def compute_fields(self, state, backend, field=("density", "velocity")):
        eng = backend.engine
        src = eng.materialize()
        from pylabolt_amd import _native as nt
        nt.check(eng.lib.d2q9_moments(eng._Lp, nt.ptr(eng.f[src]),
                                      nt.ptr(eng.mom[eng.i_mom]),
                                      eng._stream_ptr), "d2q9_moments")
        eng.mom_full = True


class StreamingOperator:
    def __init__(self, model, state, backend=None, verbose=True):
        self.streaming_type = model.streaming_type["fluid"]

    def set_backend(self, state, backend, verbose=True):
        return None

    def compile(self, state, backend, verbose=True):
        return None

    def stream(self, state, backend, fluid=True):
        """Pull streaming + boundary rules into pop_fluid_new."""
        if fluid:
            backend.engine.materialize()


class BoundaryOperator:
    def __init__(self, model, state, backend=None, verbose=True):
        self.no_of_boundaries = len(state.boundary.boundary_elements)
        self.model = model

    def set_backend(self, state, backend, verbose=True):
        return None

    def compile(self, state, backend, verbose=True):
        return None

    def set_boundary(self, state, backend, fluid=True):
        """Applied inside the stream sweep, in element order."""
        return None

    def compute_force(self, state, backend, mpi_operator, time_step=0):
        """Force on every wall element by momentum exchange, at the steps
        the reference computes it (boundary_operator.py:204-240): when
        ``compute_force`` is on and ``time_step`` is a multiple of the
        boundary write interval.  Sets ``element.force``."""
        b = state.boundary
        if not b.compute_force or b.write_interval is None:
            return
        if time_step % b.write_interval != 0:
            return
        eng = backend.engine
        if getattr(self, "_groups", None) is None:
            self._groups = wall_boundary_groups(state)
            self._handle = eng.set_force_groups(self._groups)
        if not self._groups:
            return
        out = eng.momentum_exchange(self._handle)
        for g, row in zip(self._groups, out):
            g["element"].force[0] = row[0]
            g["element"].force[1] = row[1]


class MPIOperator:
    def __init__(self, comm, state, verbose=True):
        self.comm = comm
        if state.domain.mpi_size != 1:
            raise ValueError("the GPU engine runs one sub-domain per process "
                             "and this build runs one GPU: nx = ny = 1")

    def set_backend(self, state, backend, verbose=True):
        return None

    def compile(self, state, backend, verbose=True):
        return None

    def halo_exchange(self, state, backend, bool_buffers=None,
                      int_buffers=None, float_buffers=None):
        return None

    def reduce(self, local_array, operation="sum"):
        return local_array


class ObstacleOperator:
    def __init__(self, model, state, backend=None, mpi_operator=None,
                 force_operator=None, verbose=True):
        self.no_of_obstacles = state.obstacle.no_of_obstacles
        mark_obstacle_boundaries(state)

    def set_backend(self, state, backend, verbose=True):
        return None

    def compile(self, state, backend, verbose=True):
        return None

    def move_obstacles(self, state, backend, mpi_operator):
        return None

    def compute_force_torque(self, state, backend, mpi_operator):
        """Force and torque on every obstacle by momentum exchange
        (obstacle_operator.py:75-113, force_torque_kernels.py:11-95); the
        torque is taken about the obstacle's reference point (a circle's
        centre).  Sets ``obstacle.force`` and ``obstacle.torque``."""
        obs = state.obstacle
        if not obs.compute_force_torque or obs.no_of_obstacles == 0:
            return
        eng = backend.engine
        if getattr(self, "_handle", None) is None:
            nodes, masks, ids = obstacle_boundary_nodes(state)
            groups = []
            for o in obs.obstacles:
                sel = ids == o.id
                groups.append({"nodes": nodes[sel], "masks": masks[sel],
                               "ref": tuple(float(v) for v in o.ref_point),
                               "torque": True})
            self._handle = eng.set_force_groups(groups)
        out = eng.momentum_exchange(self._handle)
        for o, row in zip(obs.obstacles, out):
            o.force[0] = row[0]
            o.force[1] = row[1]
            o.torque = row[2]
