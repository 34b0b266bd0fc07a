"""Run-control parameters (reference: pylabolt/base/control.py:31-49)."""
import numpy as np

from pylabolt_amd.utils.helpers import print_log


class Control:
    """``control_dict`` = {start_time, end_time, std_out_interval,
    save_interval, checkpoint_interval, precision: "single"|"double"}.

    ``precision`` picks the state dtype; ``float_min`` is that dtype's
    machine epsilon, the regulariser in ``1/(rho + float_min)``
    (control.py:49).
    """

    def __init__(self, simulation, rank, verbose=True):
        if not hasattr(simulation, "control_dict"):
            raise ValueError("control_dict not found in simulation.py file")
        control_dict = simulation.control_dict
        print_log("-" * 80, rank, verbose=verbose)
        print_log("Setting control parameters...", rank, verbose=verbose)
        try:
            self.start_time = control_dict["start_time"]
            self.end_time = control_dict["end_time"]
            self.std_out_interval = control_dict["std_out_interval"]
            self.save_interval = control_dict["save_interval"]
            self.checkpoint_interval = control_dict["checkpoint_interval"]
            self.precision_type = control_dict["precision"]
            if self.precision_type == "single":
                self.precision = np.float32
            elif self.precision_type == "double":
                self.precision = np.float64
            else:
                raise ValueError(
                    "unsupported precision specified." +
                    "available precision (single, double)"
                )
        except KeyError as e:
            raise ValueError(e.args[0] + " missing in control_dict")
        self.float_min = np.finfo(self.precision).eps
        print_log("Setting control parameters done!", rank, verbose=verbose)
        print_log("-" * 80, rank, verbose=verbose)
