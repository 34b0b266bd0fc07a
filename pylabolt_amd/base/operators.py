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
* ``move_obstacles`` / ``compute_force_torque``: static bodies only.
"""


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
        return None


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

    def set_backend(self, state, backend, verbose=True):
        return None

    def compile(self, state, backend, verbose=True):
        return None

    def move_obstacles(self, state, backend, mpi_operator):
        return None

    def compute_force_torque(self, state, backend, mpi_operator):
        return None
