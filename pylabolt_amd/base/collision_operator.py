"""Collision operator (reference: pylabolt/base/collision_operator.py).

Reads ``collision_dict.fluid = {model, equilibrium, forcing_model}``
(collision_operator.py:44-134), checks them against the solver's model
tables, and derives ``tau = kin_visc * inv_cs_2 + 0.5`` and
``omega = 1 / tau`` (collision_operator.py:89-91).  For MRT the relaxation
rates are ``S = (1, 1, 1, 1, 1, 1, 1, omega, omega)`` in the reference's
moment basis (collision_operator.py:136-165).

``collide`` runs moments + collision from the current ``pop_fluid_new``
into ``pop_fluid`` on the GPU (the reference's density -> force -> velocity
-> collide passes, fused).  The solver's time step does not call it: the
fused sweep does this work together with streaming.
"""
import numpy as np

from pylabolt_amd.utils.helpers import print_log


class CollisionOperator:
    def __init__(self, simulation, model, state, comm=None, verbose=True):
        if not hasattr(simulation, "collision_dict"):
            raise ValueError("collision_dict not found in simulation.py file")
        self.model = model
        self.collision_dict = simulation.collision_dict
        self.read_collision_dict(state, verbose)

    def read_collision_dict(self, state, verbose=True):
        if "fluid" not in self.collision_dict:
            raise ValueError("fluid missing in collision_dict")
        fluid_dict = self.collision_dict["fluid"]
        for key in ["model", "equilibrium", "forcing_model"]:
            if key not in fluid_dict:
                raise ValueError(key + " missing in fluid: collision_dict")
        self.collision_fluid = fluid_dict["model"]
        self.equilibrium_fluid = fluid_dict["equilibrium"]
        self.forcing_fluid = fluid_dict["forcing_model"]
        if self.collision_fluid not in self.model.collision_models["fluid"]:
            raise ValueError(
                "Unsupported fluid collision model: " +
                str(self.collision_fluid) + "\nAvailable models: " +
                str(self.model.collision_models["fluid"])
            )
        if self.equilibrium_fluid not in self.model.equilibrium_models["fluid"]:
            raise ValueError(
                "Unsupported fluid equilibrium model: " +
                str(self.equilibrium_fluid) + "\nAvailable models: " +
                str(self.model.equilibrium_models["fluid"])
            )
        if self.forcing_fluid not in self.model.forcing_models["fluid"]:
            raise ValueError(
                "Unsupported fluid forcing model: " + str(self.forcing_fluid) +
                "\nAvailable models: " +
                str(self.model.forcing_models["fluid"])
            )
        self.tau_fluid = state.transport.kin_visc * state.lattice.inv_cs_2 + 0.5
        self.omega_fluid = 1 / self.tau_fluid
        if self.collision_fluid == "MRT":
            self.setup_MRT_params(state)
        self.collision_params = (self.omega_fluid,)
        print_log(f"{'Fluid collision model':<30}: {self.collision_fluid}",
                  state.domain.mpi_rank, verbose)

    def setup_MRT_params(self, state):
        """Moment basis M and rates S (collision_operator.py:136-165)."""
        self.M = np.array([
            [1, 1, 1, 1, 1, 1, 1, 1, 1],
            [-4, -1, -1, -1, -1, 2, 2, 2, 2],
            [4, -2, -2, -2, -2, 1, 1, 1, 1],
            [0, 1, 0, -1, 0, 1, -1, -1, 1],
            [0, -2, 0, 2, 0, 1, -1, -1, 1],
            [0, 0, 1, 0, -1, 1, 1, -1, -1],
            [0, 0, -2, 0, 2, 1, 1, -1, -1],
            [0, 1, -1, 1, -1, 0, 0, 0, 0],
            [0, 0, 0, 0, 0, 1, -1, 1, -1],
        ], dtype=state.control.precision)
        self.inv_M = np.linalg.inv(self.M)
        self.S = np.array([1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0,
                           self.omega_fluid, self.omega_fluid],
                          dtype=state.control.precision)

    @property
    def rates(self):
        if self.collision_fluid == "MRT":
            return [float(v) for v in self.S]
        return [float(self.omega_fluid)] * 9

    def set_backend(self, state, backend, verbose=True):
        self.engine = backend.engine

    def compile(self, state, backend, verbose=True):
        return None

    def initialize_pop(self, state, backend):
        """pop = pop_new = feq(rho0, u0) (equilibrium_kernels.py:38-78)."""
        backend.engine.initialize()

    def collide(self, state, backend, fluid=True):
        if fluid:
            eng = backend.engine
            src = eng.materialize()
            eng._collide(src, 1 - src, True)
            eng.i_pop, eng.i_pop_new = 1 - src, None
            eng.time_step += 1
            eng.mom_full = True
