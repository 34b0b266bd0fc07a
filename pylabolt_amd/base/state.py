"""Case state (reference: pylabolt/base/state.py:14-113).

Builds control -> mesh -> lattice -> domain -> transport -> fields ->
init_fields -> boundary -> obstacle from a ``simulation`` module, in the
reference's order.  A configuration error raises ``ValueError`` with the
reference's message; the reference prints ``FATAL ERROR!`` and aborts the
MPI job instead, which a caller can still do around this constructor.
"""
from pylabolt_amd.base import (boundary, control, domain, fields, lattice,
                               mesh, obstacle, transport)
from pylabolt_amd.base.init_fields import init_fields
from pylabolt_amd.utils.helpers import SingleRankComm


class State:
    def __init__(self, simulation, comm=None, mpi_rank=0, fluid=True,
                 verbose=False):
        if comm is None:
            comm = SingleRankComm()
        self.fluid = fluid
        self.phase = False
        self.control = control.Control(simulation, mpi_rank, verbose=verbose)
        self.mesh = mesh.Mesh(simulation, mpi_rank, verbose=verbose)
        self.lattice = lattice.Lattice(simulation, self.control, self.mesh,
                                       mpi_rank, verbose=verbose)
        self.domain = domain.Domain(simulation, self.mesh, comm,
                                    verbose=verbose)
        self.transport = transport.Transport(simulation, self.control,
                                             self.domain, fluid=fluid,
                                             verbose=verbose)
        self.fields = fields.Fields(self.control, self.lattice, self.domain,
                                    fluid=fluid)
        init_fields(simulation, self.control, self.domain, self.fields,
                    fluid=fluid, verbose=verbose)
        self.boundary = boundary.Boundary(simulation, self.mesh, self.domain,
                                          self.control, self.fields,
                                          fluid=fluid, verbose=verbose)
        self.obstacle = obstacle.Obstacle(simulation, self.mesh, self.domain,
                                          self.control, self.fields,
                                          self.boundary, fluid=fluid,
                                          verbose=verbose)
