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
        self.history_save_path = os.path.join(root, "output", "histories")
        self._dumped = False
        self.setup_write_histories(state)

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
        eng = getattr(self, "engine", None)
        if eng is not None:
            # not read by the reference's tools: how the run was computed
            meta["engine"] = {"storage": eng.storage_name,
                              "two_step": bool(eng.two_step)}
        os.makedirs(self.field_save_path, exist_ok=True)
        with open(os.path.join(self.root, "metadata.json"), "w") as f:
            json.dump(meta, f, indent=4)
        self._dumped = True

    def set_backend(self, state, backend, verbose=True):
        self.engine = backend.engine

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

    def setup_write_histories(self, state):
        """History files of the obstacles and wall elements, with the
        reference's headers (io_operator.py:236-296)."""
        obs, bnd = state.obstacle, state.boundary
        if not (obs.write_obstacle_data or bnd.write_boundary_data):
            return
        os.makedirs(self.history_save_path, exist_ok=True)
        if obs.write_obstacle_data:
            for o in obs.obstacles:
                with open(os.path.join(self.history_save_path,
                                       o.name + ".dat"), "w") as fh:
                    fh.write(f"{'#':5} {'PyLaBolt obstacle history'}\n"
                             f"{'#':5} {'ID':8}: {o.id}\n"
                             f"{'#':5} {'Name':8}: {o.name}\n"
                             f"{'#':5} {'Type':8}: {o.type}\n"
                             f"{'#':5} {'Columns':8}:\n")
                    fh.write(f"{'#':5}{'time':<21}{'pos_x':<30}{'pos_y':<30}"
                             f"{'alpha':<30}{'vel_x':<30}{'vel_y':<30}"
                             f"{'omega':<30}{'force_x':<30}{'force_y':<30}"
                             f"{'torque':<30}\n")
        if bnd.write_boundary_data:
            for e in bnd.boundary_elements:
                if not e.wall:
                    continue
                with open(os.path.join(self.history_save_path,
                                       e.name + ".dat"), "w") as fh:
                    fh.write(f"{'#':5} {'PyLaBolt boundary history'}\n"
                             f"{'#':5} {'Name':8}: {e.name}\n"
                             f"{'#':5} {'Columns':8}:\n")
                    fh.write(f"{'#':5}{'time':<21}{'force_x':<30}"
                             f"{'force_y':<30}\n")

    def write_histories(self, state, time_step):
        """Append one line per obstacle / wall element at the write
        interval (io_operator.py:297-357)."""
        obs, bnd = state.obstacle, state.boundary
        if (obs.write_obstacle_data and obs.write_interval is not None
                and time_step % obs.write_interval == 0):
            for o in obs.obstacles:
                with open(os.path.join(self.history_save_path,
                                       o.name + ".dat"), "a") as fh:
                    fh.write(f"{time_step:<24}"
                             f"{o.center[0]:<30.16e}{o.center[1]:<30.16e}"
                             f"{o.inclination_angle:<30.16e}"
                             f"{o.linear_velocity[0]:<30.16e}"
                             f"{o.linear_velocity[1]:<30.16e}"
                             f"{o.angular_velocity:<30.16e}"
                             f"{o.force[0]:<30.16e}{o.force[1]:<30.16e}"
                             f"{o.torque:<30.16e}\n")
        if (bnd.write_boundary_data and bnd.write_interval is not None
                and time_step % bnd.write_interval == 0):
            for e in bnd.boundary_elements:
                if not e.wall:
                    continue
                with open(os.path.join(self.history_save_path,
                                       e.name + ".dat"), "a") as fh:
                    fh.write(f"{time_step:<24}{e.force[0]:<30.16e}"
                             f"{e.force[1]:<30.16e}\n")
