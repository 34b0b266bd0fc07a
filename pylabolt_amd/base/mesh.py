"""Global grid description (reference: pylabolt/base/mesh.py:20-40)."""
import numpy as np

from pylabolt_amd.utils.helpers import print_log


class Mesh:
    """``mesh_dict = {grid: [Nx, Ny]}``; validation messages as the
    reference's."""

    def __init__(self, simulation, rank, verbose=True):
        if not hasattr(simulation, "mesh_dict"):
            raise ValueError("mesh_dict not found in simulation.py file")
        mesh_dict = simulation.mesh_dict
        if "grid" not in mesh_dict:
            raise ValueError("grid missing in mesh_dict")
        grid = mesh_dict["grid"]
        if not isinstance(grid, list) or len(grid) != 2:
            raise ValueError("grid entry in mesh_dict must be a list [Nx, Ny]")
        grid = np.array(grid, dtype=int)
        if np.any(grid == 0):
            raise ValueError("grid dimensions cannot be zero")
        if np.all(grid == 1):
            raise ValueError("grid specified is a point")
        if grid[0] == 1 or grid[1] == 1:
            self.dimensions = 1
        else:
            self.dimensions = 2
        self.grid_global_shape = grid
        self.grid_global_size = np.prod(self.grid_global_shape)
        print_log(
            "global grid size set: (" + str(self.grid_global_shape) + ")",
            rank, verbose
        )
