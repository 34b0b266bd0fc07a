"""Transport properties (reference: pylabolt/base/transport.py:20-33)."""
from pylabolt_amd.utils.helpers import print_log


class Transport:
    """``transport_dict = {kin_visc}``; the relaxation time follows as
    ``tau = kin_visc * inv_cs_2 + 0.5`` in the collision operator."""

    def __init__(self, simulation, control, domain, fluid=True, verbose=True):
        print_log("-" * 80, domain.mpi_rank, verbose)
        print_log("Setting transport properties...\n", domain.mpi_rank, verbose)
        if not hasattr(simulation, "transport_dict"):
            raise ValueError("transport_dict not found in simulation.py file")
        transport_dict = simulation.transport_dict
        if "kin_visc" not in transport_dict:
            raise ValueError("kin_visc missing in transport_dict")
        self.kin_visc = transport_dict["kin_visc"]
        if not isinstance(self.kin_visc, (float, int)):
            raise ValueError("kin_visc must be a float/int")
        self.kin_visc = control.precision(self.kin_visc)
        print_log(f"{'kinematic viscosity':20s}: {self.kin_visc}",
                  domain.mpi_rank, verbose)
        print_log("Setting transport properties done!", domain.mpi_rank, verbose)
        print_log("-" * 80, domain.mpi_rank, verbose)
