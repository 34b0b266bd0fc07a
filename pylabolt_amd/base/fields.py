"""Host-side field containers (reference: pylabolt/base/fields.py:50-92,
166-179).

These are the reference's arrays with the reference's shapes and dtypes:
AoS populations ``pop_fluid[size, 9]``, ``velocity[size, 2]`` and masks over
the padded grid.  They describe the case and receive results copied back
from the GPU; the time step itself runs on the engine's device buffers
(structure-of-arrays, pitched), never on these.
"""
import numpy as np


class Fields:
    def __init__(self, control, lattice, domain, fluid=True):
        self.fluid = fluid
        size = domain.size
        prec = control.precision
        self.solid = np.zeros(size, dtype=np.bool_)
        self.solid_id = np.full(size, -1, dtype=int)
        self.solid_boundary = np.zeros(size, dtype=np.bool_)
        self.fluid_boundary = np.zeros(size, dtype=np.bool_)
        self.surface_normals = np.zeros((size, 2), dtype=prec)
        self.ghost_node = np.zeros(size, dtype=np.bool_)
        self.init_ghost_nodes(domain)
        self.periodic_boundary = np.zeros(size, dtype=np.bool_)
        if self.fluid:
            self.velocity = np.zeros((size, 2), dtype=prec)
            self.density = np.zeros(size, dtype=prec)
            self.pressure = np.zeros(size, dtype=prec)
            self.force_field = np.zeros((size, 2), dtype=prec)
            self.pop_fluid = np.zeros((size, lattice.no_of_directions),
                                      dtype=prec)
            self.pop_fluid_new = np.zeros((size, lattice.no_of_directions),
                                          dtype=prec)

    def init_ghost_nodes(self, domain):
        """Mark the one-cell ring ``i in {0, Nx_pad-1}`` or
        ``j in {0, Ny_pad-1}`` (fields.py:166-179)."""
        Nx, Ny = (int(v) for v in domain.shape)
        g = self.ghost_node.reshape(Nx, Ny)
        g[0, :] = True
        g[-1, :] = True
        g[:, 0] = True
        g[:, -1] = True
