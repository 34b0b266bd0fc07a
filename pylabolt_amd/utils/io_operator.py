"""Field output (reference: pylabolt/utils/io_operator.py:41-234 and
parallel/cpu/io_operator_kernels.py:5-52).

Every ``save_interval`` steps the inner domain of each saved field is
written to ``output/fields/t_<step>.npz`` (inner index
``(i - 1) * Ny + (j - 1)``), and ``metadata.json`` describes the run, in
the reference's format so its ``reconstruct``/``npz2vtk`` tools read it.
"""
import json
import os

import numpy as np


class InputOutputOperator:
    def __init__(self, model, state, backend=None, comm=None, verbose=True,
                 root="."):
        self.model = model
        self.root = root
        self.fields_list = model.save_fields
        self.fields_save_metadata = {}
        for name in self.fields_list:
            if not hasattr(state.fields, name):
                raise ValueError(name + " is not a valid field for saving")
            field = getattr(state.fields, name)
            comps = 1 if field.ndim == 1 else field.shape[1]
            self.fields_save_metadata[name] = {"components": comps,
                                               "dtype": str(field.dtype)}
        self.field_save_path = os.path.join(root, "output", "fields")
        self._dumped = False

    def dump_metadata(self, state):
        meta = {
            "pylabolt": {"version": "pylabolt_amd", "solver":
                         self.model.solver_name},
            "control": {"end_time": state.control.end_time,
                        "start_time": state.control.start_time,
                        "save_interval": state.control.save_interval,
                        "checkpoint_interval":
                            state.control.checkpoint_interval},
            "mesh": {"size": int(state.mesh.grid_global_size),
                     "shape": (int(state.mesh.grid_global_shape[0]),
                               int(state.mesh.grid_global_shape[1]))},
            "decomposition": {"nx": int(state.domain.no_of_procs_x),
                              "ny": int(state.domain.no_of_procs_y)},
            "fields_saved": self.fields_save_metadata,
        }
        os.makedirs(self.field_save_path, exist_ok=True)
        with open(os.path.join(self.root, "metadata.json"), "w") as f:
            json.dump(meta, f, indent=4)
        self._dumped = True

    def set_backend(self, state, backend, verbose=True):
        return None

    def compile(self, state, backend, verbose=True):
        return None

    @staticmethod
    def inner(state, field):
        nxp, nyp = (int(v) for v in state.domain.shape)
        if field.ndim == 1:
            return np.ascontiguousarray(field.reshape(nxp, nyp)[1:-1, 1:-1]
                                        .reshape(-1))
        c = field.shape[1]
        return np.ascontiguousarray(
            field.reshape(nxp, nyp, c)[1:-1, 1:-1, :].reshape(-1, c))

    def write_fields(self, state, backend, time_step, solver=None):
        if state.control.save_interval is None:
            return
        if time_step % state.control.save_interval != 0:
            return
        if not self._dumped:
            self.dump_metadata(state)
        if solver is not None:
            solver.sync_fields()
        out = {name: self.inner(state, getattr(state.fields, name))
               for name in self.fields_list}
        np.savez(os.path.join(self.field_save_path,
                              "t_" + str(time_step) + ".npz"), **out)

    def write_histories(self, state, time_step):
        return None
