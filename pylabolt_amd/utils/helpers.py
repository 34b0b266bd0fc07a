"""Logging and case-loading helpers.

Mirrors pylabolt/utils/helpers.py:6-56 of the reference: ``print_log`` prints
on rank 0 only, ``load_simulation`` imports the user's ``simulation.py`` from
the current working directory, and ``SimulationStatusLogger`` prints the
``time | res_density | res_velocity`` line every ``std_out_interval`` steps.
"""
import importlib
import os
import sys

import numpy as np


def print_log(mssg, mpi_rank, verbose):
    """Print ``mssg`` on rank 0 when ``verbose`` (helpers.py:6-8)."""
    if verbose and mpi_rank == 0:
        print(mssg, flush=True)


class SingleRankComm:
    """The communicator of a one-GPU run: rank 0 of a world of size 1.

    Stands where the reference passes ``mpi4py.MPI.COMM_WORLD``; only the
    calls the setup path makes are provided.  ``Abort`` raises instead of
    killing the process so that callers (and tests) see the error.
    """

    def Get_rank(self):
        return 0

    def Get_size(self):
        return 1

    def Barrier(self):
        return None

    def Abort(self, errorcode=1):
        raise SystemExit(errorcode)


def load_simulation(comm=None, mpi_rank=0, path=None):
    """Import ``simulation.py`` from ``path`` (default: the cwd).

    Follows helpers.py:11-26: the case file is a Python module whose
    ``*_dict`` attributes describe the run.  A missing file raises
    ``ImportError`` with the reference's message.
    """
    working_dir = os.getcwd() if path is None else str(path)
    candidate = os.path.join(working_dir, "simulation.py")
    if not os.path.isfile(candidate):
        raise ImportError(
            "Missing simulation.py file in current working directory"
        )
    spec = importlib.util.spec_from_file_location("simulation", candidate)
    module = importlib.util.module_from_spec(spec)
    sys.modules["simulation"] = module
    spec.loader.exec_module(module)
    return module


class SimulationStatusLogger:
    """Status line printer (helpers.py:29-56)."""

    def __init__(self, mpi_rank, verbose=True):
        self.verbose = verbose
        self.lines = []

    def log_data(self, state, time_step, **values):
        if state.control.std_out_interval is None:
            return
        if time_step % state.control.std_out_interval != 0:
            return
        log_string = [f"{'time:':<5} {time_step:<10}"]
        for key, value in values.items():
            if np.isscalar(value):
                value_str = f"{value:.5e}"
            elif len(value) == 1:
                value_str = f"{value[0]:.5e}"
            else:
                value_str = "(" + ", ".join(f"{v:.5e}" for v in value) + ")"
            log_string.append(f"{key:<5}: {value_str}")
        line = " | ".join(log_string)
        self.lines.append(line)
        print_log(line, state.domain.mpi_rank, self.verbose)
