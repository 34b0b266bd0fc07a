"""Sub-domain description (reference: pylabolt/parallel/domain.py:5-85).

The padded shape ``(Nx_rank + 2, Ny_rank + 2)`` with a one-cell ghost ring
and the flat index ``ind = i * Ny_pad + j`` (y fastest) are the reference's
conventions; every host array and every AoS view this package exposes uses
them.  The block split across ranks is kept as host bookkeeping: the engine
runs one rank per GPU and this build runs one GPU.
"""
import numpy as np


class Domain:
    def __init__(self, simulation, mesh, comm, verbose=True):
        self.mpi_rank = comm.Get_rank()
        self.mpi_size = comm.Get_size()
        if not hasattr(simulation, "decompose_dict"):
            raise ValueError("decompose_dict not found in simulation.py file")
        decompose_dict = simulation.decompose_dict
        if ("nx" not in decompose_dict or "ny" not in decompose_dict):
            raise ValueError("nx or ny missing decompose_dict")
        self.no_of_procs_x = decompose_dict["nx"]
        self.no_of_procs_y = decompose_dict["ny"]
        if self.mpi_size != self.no_of_procs_x * self.no_of_procs_y:
            raise ValueError(
                "invalid domain decomposition. " +
                "nx * ny not equal to total no.of MPI processes"
            )
        self.i_proc = self.mpi_rank // self.no_of_procs_y
        self.j_proc = self.mpi_rank % self.no_of_procs_y
        gx, gy = (int(v) for v in mesh.grid_global_shape)
        bx = int(np.ceil(gx / self.no_of_procs_x))
        by = int(np.ceil(gy / self.no_of_procs_y))
        if self.i_proc != self.no_of_procs_x - 1:
            self.Nx_rank = bx
        else:
            self.Nx_rank = int(gx - self.i_proc * bx)
        offset_x = self.i_proc * bx
        if self.j_proc != self.no_of_procs_y - 1:
            self.Ny_rank = by
        else:
            self.Ny_rank = int(gy - self.j_proc * by)
        offset_y = self.j_proc * by
        self.offset = np.array([offset_x, offset_y], dtype=int)
        self.Nx_pad = self.Nx_rank + 2
        self.Ny_pad = self.Ny_rank + 2
        self.shape = np.array([self.Nx_pad, self.Ny_pad])
        self.size = int(np.prod(self.shape))
        self.inner_shape = np.array([self.Nx_rank, self.Ny_rank])
        self.inner_size = int(np.prod(self.inner_shape))


def local_to_global(i_local, j_local, offset):
    """Reference parallel/cpu/MPI_kernels.py local_to_global: add offset."""
    return i_local + offset[0], j_local + offset[1]


def global_to_local(i_global, j_global, offset):
    """Reference parallel/cpu/MPI_kernels.py global_to_local."""
    return i_global - offset[0], j_global - offset[1]
