"""Force operator (reference: pylabolt/base/force_operator.py:46-79).

``forcing_dict = {gravity: [gx, gy]}``; gravity is ignored (zero) when the
collision dict's ``forcing_model`` is None.  The force field ``rho * g`` is
computed inside the fused sweep, so ``compute_force_field`` has nothing to
launch; ``force_field_host`` reproduces the reference array for output.
"""
import numpy as np

from pylabolt_amd.utils.helpers import print_log


class ForceOperator:
    def __init__(self, simulation, model, state, comm=None,
                 collision_operator=None, verbose=True):
        if not hasattr(simulation, "forcing_dict"):
            raise ValueError("forcing_dict not found in simulation.py file")
        self.forcing_dict = simulation.forcing_dict
        self.collision_operator = collision_operator
        self.gravity = np.zeros(2, dtype=state.control.precision)
        if "gravity" in self.forcing_dict:
            if collision_operator.forcing_fluid is None:
                print_log(
                    "WARNING! gravity ignored in forcing dict" +
                    " as forcing is set to None in collision dict\n",
                    state.domain.mpi_rank, verbose)
            else:
                gravity = self.forcing_dict["gravity"]
                if not isinstance(gravity, list) or len(gravity) != 2:
                    raise ValueError(
                        "gravity must be a list (gx, gy) in forcing dict")
                self.gravity = np.array(gravity, dtype=state.control.precision)

    def set_backend(self, state, backend, verbose=True):
        return None

    def compile(self, state, backend, verbose=True):
        return None

    def compute_force_field(self, state, backend):
        """Computed inside the sweep (force_field_kernels.py:5-25)."""
        return None

    def force_field_host(self, state):
        f = np.zeros((state.domain.size, 2), dtype=np.float64)
        m = ~state.fields.solid & ~state.fields.ghost_node
        f[m, 0] = state.fields.density[m] * self.gravity[0]
        f[m, 1] = state.fields.density[m] * self.gravity[1]
        return f
