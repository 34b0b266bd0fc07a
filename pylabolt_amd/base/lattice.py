"""D2Q9 lattice constants (reference: pylabolt/base/lattice.py:41-60).

The constants are computed exactly the way the reference computes them, so
that ``inv_cs_2`` is 2.999999999999999 and ``inv_cs_4`` 8.999999999999995 in
double precision (not 3 and 9): every kernel takes them from here.
"""
import numpy as np

from pylabolt_amd.utils.helpers import print_log

# D2Q9 velocity set in the reference's ordering (lattice.py:50-59).
CX = (0, 1, 0, -1, 0, 1, -1, -1, 1)
CY = (0, 0, 1, 0, -1, 1, 1, -1, -1)
INV = (0, 3, 4, 1, 2, 7, 8, 5, 6)


class Lattice:
    def __init__(self, simulation, control, mesh, rank, verbose=True):
        if not hasattr(simulation, "lattice_dict"):
            raise ValueError("lattice_dict not found in simulation.py file")
        lattice_dict = simulation.lattice_dict
        if "lattice_type" not in lattice_dict:
            raise ValueError("lattice_type missing in lattice_dict")
        self.lattice_type = lattice_dict["lattice_type"]
        self.cs = control.precision(1/np.sqrt(3))
        self.cs_2 = (self.cs * self.cs)
        self.inv_cs_2 = 1.0 / self.cs_2
        self.inv_cs_4 = self.inv_cs_2 * self.inv_cs_2
        if self.lattice_type == "D2Q9":
            if mesh.dimensions != 2:
                raise ValueError(
                    "grid dimensions and lattice type are incompatible"
                )
            self.cx = np.array(CX, dtype=int)
            self.cy = np.array(CY, dtype=int)
            self.weights = np.array(
                [4/9, 1/9, 1/9, 1/9, 1/9, 1/36, 1/36, 1/36, 1/36],
                dtype=control.precision
            )
            self.inv_list = np.array(INV, dtype=int)
            self.no_of_directions = int(9)
        else:
            # D1Q3 exists in the reference (lattice.py:61-70) but the fluidLB
            # time step this engine implements is D2Q9 only.
            raise ValueError("Unsupported lattice type")
        print_log("lattice type set: " + self.lattice_type, rank, verbose)
