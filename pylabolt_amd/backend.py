"""Runtime backend (reference: pylabolt/parallel/backend.py:19-131).

The reference picks Numba CPU threads or one CUDA device with a CuPy stream
and graph.  Here the only backend is one MI355X (gfx950) per process: the
D2Q9 engine with its own HIP stream and HIP graphs.  Asking for the "cpu"
backend is an error rather than a silent fallback.
"""
import torch

from pylabolt_amd import _native as nt
from pylabolt_amd.engine import D2Q9Engine, elements_from_boundary


class Backend:
    def __init__(self, comm, state, backend_type="gpu", n_threads=None,
                 storage="f64", tile=(64, 4), device=None, graph_steps=16,
                 stream=None, solid_tiles=None, two_step=None):
        if backend_type != "gpu":
            raise ValueError("Unsupported backend: " + str(backend_type) +
                             " (this engine runs on the GPU only)")
        if state.domain.mpi_size > 1:
            raise ValueError("GPU backend runs one sub-domain per process")
        self.backend_type = "gpu"
        self.storage = storage
        self.tile = tile
        self.graph_steps = graph_steps
        self.stream = stream
        self.solid_tiles = solid_tiles
        self.two_step = two_step
        self.lib = nt.lib()
        self.device = torch.device(device) if device else torch.device(
            "cuda", torch.cuda.current_device())
        self.engine = None

    def create_engine(self, state, collision_operator, force_operator,
                      force_field=None, dev_scale=None):
        lat = state.lattice
        self.engine = D2Q9Engine(
            int(state.domain.Nx_rank), int(state.domain.Ny_rank),
            state.fields.solid, state.fields.density, state.fields.velocity,
            elements_from_boundary(state.boundary),
            state.boundary.x_periodic, state.boundary.y_periodic,
            collision_operator.collision_fluid,
            collision_operator.forcing_fluid, collision_operator.omega_fluid,
            force_operator.gravity, lat.inv_cs_2, lat.inv_cs_4, lat.weights,
            state.control.float_min, storage=self.storage,
            force_field=force_field, tile=self.tile,
            rates=collision_operator.rates, dev_scale=dev_scale,
            device=self.device, graph_steps=self.graph_steps,
            stream=self.stream, solid_tiles=self.solid_tiles,
            two_step=self.two_step)
        return self.engine
