"""Static obstacles (reference: pylabolt/base/obstacle.py:7-567 and
pylabolt/parallel/cpu/obstacle_kernels.py:11-93).

``obstacle_dict = {options: {...}, <name>: {type: circle, radius, center,
density, static}}``.  A non-ghost node is solid when its squared distance
from the centre (minimum image across periodic sides) is ``<= radius**2``,
with node coordinates ``x = i - 1 + offset_x``, ``y = j - 1 + offset_y``.
Solid nodes get ``solid_id``, the obstacle's density and the rigid-body
velocity ``(U - Omega * r_y, V + Omega * r_x)`` (obstacle_kernels.py:80-93).

Only static bodies are supported: the reference's moving-body path calls a
method it never defines (obstacle_operator.py:67), and no workload here
moves a body.
"""
import numpy as np

from pylabolt_amd.utils.helpers import print_log


class Circle:
    def __init__(self, obstacle_id, name, user_dict, control, domain, mesh,
                 fields, boundary):
        self.id = obstacle_id
        self.name = name
        self.type = "circle"
        self.force = np.zeros(2, dtype=control.precision)
        self.torque = control.precision(0)
        self.linear_velocity = np.zeros(2, dtype=control.precision)
        self.angular_velocity = control.precision(0)
        if "radius" not in user_dict:
            raise ValueError("radius missing in obstacle: " + name)
        radius = user_dict["radius"]
        if type(radius) not in (float, int):
            raise ValueError(
                "radius must be float or int for obstacle type circle: " + name
            )
        self.radius = control.precision(radius)
        if "center" not in user_dict:
            raise ValueError("center missing in obstacle: " + name)
        center = user_dict["center"]
        if not isinstance(center, list):
            raise ValueError(
                "center must be list for obstacle type circle: " + name
            )
        self.center = np.array(center, dtype=control.precision)
        # torque reference point of a circle: its centre (obstacle.py:450)
        self.ref_point = self.center
        self.inclination_angle = control.precision(0)
        if "density" not in user_dict:
            raise ValueError("density missing in obstacle: " + name)
        solid_density = user_dict["density"]
        if type(solid_density) not in (float, int):
            raise ValueError(
                "density must be float or int for obstacle type circle: " +
                name
            )
        self.solid_density = control.precision(solid_density)
        if "static" not in user_dict:
            raise ValueError("static missing in obstacle: " + name)
        static = user_dict["static"]
        if not isinstance(static, (bool, np.bool_)):
            raise ValueError(
                "static must be True/False for obstacle type circle: " + name
            )
        self.static = static
        if not self.static:
            raise ValueError(
                "moving obstacles are not supported: " + name
            )
        construct_circle(domain, mesh, boundary, fields, self.center,
                         self.radius, self.solid_density,
                         self.linear_velocity, self.angular_velocity, self.id)


def construct_circle(domain, mesh, boundary, fields, center, radius,
                     solid_density, linear_velocity, angular_velocity,
                     solid_id):
    """Rasterise one circle into the solid mask (obstacle_kernels.py:11-93),
    vectorised over the sub-domain."""
    nxp, nyp = (int(v) for v in domain.shape)
    gx, gy = (int(v) for v in mesh.grid_global_shape)
    ii, jj = np.meshgrid(np.arange(nxp), np.arange(nyp), indexing="ij")
    x_global = (ii - 1 + int(domain.offset[0])).ravel()
    y_global = (jj - 1 + int(domain.offset[1])).ravel()
    rx = x_global - center[0]
    ry = y_global - center[1]
    rx_min = rx.copy()
    ry_min = ry.copy()
    if boundary.x_periodic:
        sel = np.abs(rx + gx) < np.abs(rx_min)
        rx_min[sel] = (rx + gx)[sel]
        sel = np.abs(rx - gx) < np.abs(rx_min)
        rx_min[sel] = (rx - gx)[sel]
    if boundary.y_periodic:
        sel = np.abs(ry + gy) < np.abs(ry_min)
        ry_min[sel] = (ry + gy)[sel]
        sel = np.abs(ry - gy) < np.abs(ry_min)
        ry_min[sel] = (ry - gy)[sel]
    dist_sq = rx_min * rx_min + ry_min * ry_min
    inside = (dist_sq <= radius * radius) & ~fields.ghost_node
    fields.solid[inside] = True
    fields.solid_id[inside] = solid_id
    fields.velocity[inside, 0] = (linear_velocity[0] -
                                  angular_velocity * ry_min[inside])
    fields.velocity[inside, 1] = (linear_velocity[1] +
                                  angular_velocity * rx_min[inside])
    fields.density[inside] = solid_density


class Obstacle:
    def __init__(self, simulation, mesh, domain, control, fields, boundary,
                 fluid=True, verbose=True):
        print_log("-" * 80, domain.mpi_rank, verbose)
        print_log("Setting up obstacles...\n", domain.mpi_rank, verbose)
        if not hasattr(simulation, "obstacle_dict"):
            raise ValueError("obstacle_dict not found in simulation.py file")
        self.obstacle_dict = simulation.obstacle_dict
        self.compute_force_torque = False
        self.ref_point_torque = np.zeros(2, dtype=control.precision)
        self.write_obstacle_data = False
        self.write_interval = 1
        self.obstacles = []
        if "options" not in self.obstacle_dict:
            raise ValueError("options missing in obstacle_dict")
        options = self.obstacle_dict["options"]
        if "compute_force_torque" in options:
            self.compute_force_torque = options["compute_force_torque"]
            if not isinstance(self.compute_force_torque, (bool, np.bool_)):
                raise ValueError(
                    "compute_force_torque must be bool:" +
                    " True/False (default: False)"
                )
        # options of obstacle.py:147-190
        if "ref_point_torque" in options:
            if self.compute_force_torque is False:
                print_log("WARNING! ref_point_torque ignored in "
                          "obstacle_dict options as compute_force_torque "
                          "is False", domain.mpi_rank, verbose)
            else:
                ref_point_torque = options["ref_point_torque"]
                if not isinstance(ref_point_torque, list):
                    raise ValueError("ref_point_torque must be list: (x, y)")
                self.ref_point_torque = np.array(ref_point_torque,
                                                 dtype=control.precision)
        if "write_obstacle_data" in options:
            self.write_obstacle_data = True
            temp_dict = options["write_obstacle_data"]
            if "interval" not in temp_dict:
                raise ValueError("interval missing obstacle_dict options!")
            self.write_interval = temp_dict["interval"]
            if not isinstance(self.write_interval, (int, type(None))):
                raise ValueError(
                    "interval must be int or None in obstacle_dict options")
            if isinstance(self.write_interval, int) and self.write_interval <= 0:
                raise ValueError("if interval is int then it must be > 0"
                                 " in obstacle_dict options")
        keys = [k for k in self.obstacle_dict.keys() if k != "options"]
        for obstacle_id, key in enumerate(keys):
            user_dict = self.obstacle_dict[key]
            if "type" not in user_dict:
                raise ValueError("type missing in obstacle: " + key)
            if user_dict["type"] == "circle":
                self.obstacles.append(Circle(obstacle_id, key, user_dict,
                                             control, domain, mesh, fields,
                                             boundary))
            else:
                raise ValueError("Unsupported obstacle type: " +
                                 str(user_dict["type"]))
        self.no_of_obstacles = len(self.obstacles)
        self.all_obstacles_static = True
        print_log("Setting up obstacles done!", domain.mpi_rank, verbose)
        print_log("-" * 80, domain.mpi_rank, verbose)
