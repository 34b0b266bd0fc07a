"""Domain boundaries (reference: pylabolt/base/boundary.py:7-603).

``boundary_dict = {options: {...}, <name>: {wall, segments, fluid: {type,
value}}}``.  Each segment of each named entry becomes one
``BoundaryElement`` in dict order, then segment order (boundary.py:474-508);
that order is the order the boundary conditions are applied in, so the last
element that writes a direction on a node wins (corners of a cavity whose
lid is listed last take the lid's moving-wall value).

Per side, ``out_list`` are the directions leaving the domain through the
wall and ``inv_list`` their reverses, the unknown incoming directions
(boundary.py:63-78).  Boundary nodes are the first interior row/column next
to the wall, flat index ``(i + 1) * Ny_pad + (j + 1)`` for global ``(i, j)``
on the segment (boundary.py:86-127).

Supported fluid types are the reference's ``bounce_back``,
``fixed_velocity``, ``fixed_pressure`` and ``periodic``, plus
``zero_gradient``.  The reference names zero-gradient in its docs but its
validator rejects it and it has no kernel; here it copies the unknown
directions from the inward neighbour after streaming.  The reference's
``fixed_pressure`` kernel indexes with the float normal
(fluid_boundary_kernels.py:133-135); here ``normal`` is the same vector as
integers, which is what that kernel means.
"""
import numpy as np

from pylabolt_amd.base.domain import global_to_local
from pylabolt_amd.utils.helpers import print_log

SUPPORTED_FLUID_BCS = [
    "bounce_back",
    "fixed_velocity",
    "fixed_pressure",
    "periodic",
    "zero_gradient",
]

# side -> (inward normal, out_list, inv_list), boundary.py:63-78
_SIDES = {
    "bottom": ((0, 1), (4, 7, 8), (2, 5, 6)),
    "top": ((0, -1), (2, 5, 6), (4, 7, 8)),
    "left": ((1, 0), (3, 6, 7), (1, 8, 5)),
    "right": ((-1, 0), (1, 8, 5), (3, 6, 7)),
}


class BoundaryElement:
    def __init__(self, boundary_name, segment, segment_no, orientation,
                 location, domain, control, fields, wall=False,
                 fluid_config=None):
        self.name = boundary_name + "_" + str(int(segment_no))
        self.orientation = orientation
        self.location = location
        self.segment = np.array(segment)
        self.periodic = False
        self.wall = wall
        self.force = np.zeros(2, dtype=control.precision)
        self.type_fluid = fluid_config["type"]
        if self.type_fluid == "periodic":
            self.periodic = True
        self.scalar_fluid = control.precision(0)
        self.vector_fluid = np.zeros(2, dtype=control.precision)
        if fluid_config["scalar_value"] is not None:
            self.scalar_fluid = control.precision(fluid_config["scalar_value"])
        if fluid_config["vector_value"] is not None:
            self.vector_fluid = np.array(fluid_config["vector_value"],
                                         dtype=control.precision)
        normal, out_list, inv_list = _SIDES[self.location]
        self.surface_normals = np.array(normal, dtype=control.precision)
        self.normal = np.array(normal, dtype=int)
        self.out_list = np.array(out_list, dtype=int)
        self.inv_list = np.array(inv_list, dtype=int)
        self.boundary_nodes = self.allocate_boundary_nodes(segment, domain)
        if self.periodic:
            fields.periodic_boundary[self.boundary_nodes] = True

    def allocate_boundary_nodes(self, segment, domain):
        """Flat indices of the segment's nodes owned by this sub-domain
        (boundary.py:86-127)."""
        nodes = []
        (x_min, y_min), (x_max, y_max) = segment
        if self.orientation == "horizontal":
            points = [(ig, y_min) for ig in range(x_min, x_max + 1)]
        else:
            points = [(x_min, jg) for jg in range(y_min, y_max + 1)]
        nx_in, ny_in = int(domain.shape[0]) - 2, int(domain.shape[1]) - 2
        for ig, jg in points:
            i, j = global_to_local(ig, jg, domain.offset)
            if i < 0 or j < 0 or i >= nx_in or j >= ny_in:
                continue
            nodes.append((i + 1) * int(domain.shape[1]) + (j + 1))
        return np.array(nodes, dtype=int)


class Boundary:
    def __init__(self, simulation, mesh, domain, control, fields, fluid=True,
                 verbose=True):
        print_log("-" * 80, domain.mpi_rank, verbose)
        print_log("Setting up domain boundaries...\n", domain.mpi_rank, verbose)
        if not hasattr(simulation, "boundary_dict"):
            raise ValueError("boundary_dict not found in simulation.py file")
        self.fluid = fluid
        self.boundary_dict = simulation.boundary_dict
        self.compute_force = False
        self.write_boundary_data = False
        self.write_interval = 1
        self.x_periodic = False
        self.y_periodic = False
        self.boundary_elements = []
        self.read_boundary_dict(mesh, domain, control, fields, verbose)
        self.no_of_boundaries = len(self.boundary_elements)
        print_log("Setting up domain boundaries done!", domain.mpi_rank,
                  verbose)
        print_log("-" * 80, domain.mpi_rank, verbose)

    def read_boundary_dict(self, mesh, domain, control, fields, verbose):
        if "options" not in self.boundary_dict:
            raise ValueError("options missing in boundary_dict")
        self.read_options_dict(self.boundary_dict["options"])
        for key in self.boundary_dict.keys():
            if key == "options":
                continue
            self.read_user_boundary_dict(key, self.boundary_dict[key], mesh,
                                         domain, control, fields, verbose)

    def read_options_dict(self, options_dict):
        """boundary.py:228-281."""
        if "compute_force" in options_dict:
            self.compute_force = options_dict["compute_force"]
            if not isinstance(self.compute_force, (bool, np.bool_)):
                raise ValueError(
                    "compute_force must be a bool:" +
                    " True/False (default: False)"
                )
        if "write_boundary_data" in options_dict:
            self.write_boundary_data = True
            temp_dict = options_dict["write_boundary_data"]
            if "interval" not in temp_dict:
                raise ValueError("interval missing boundary_dict options!")
            self.write_interval = temp_dict["interval"]
            if not isinstance(self.write_interval, (int, type(None))):
                raise ValueError(
                    "interval must be int or None in boundary_dict options"
                )
            if isinstance(self.write_interval, int) and self.write_interval <= 0:
                raise ValueError(
                    "if interval is int then it must be > 0"
                    " in boundary_dict options"
                )

    def read_user_boundary_dict(self, name, user_dict, mesh, domain, control,
                                fields, verbose):
        """Validate one named entry and create its elements
        (boundary.py:283-519)."""
        fluid_config = {"type": None, "scalar_value": None,
                        "vector_value": None}
        if "wall" not in user_dict:
            raise ValueError("wall missing in boundary: " + name)
        wall = user_dict["wall"]
        if not isinstance(wall, (bool, np.bool_)):
            raise ValueError("wall entry must be True/False: " + name)
        if "segments" not in user_dict:
            raise ValueError("segments missing in boundary: " + name)
        segments = user_dict["segments"]
        if not isinstance(segments, list):
            raise ValueError("segments must be a list object: " + name)
        if len(segments) == 0:
            raise ValueError("segments cannot be an empty list: " + name)
        for segment in segments:
            if not (isinstance(segment, list) and len(segment) == 2 and
                    all(isinstance(pt, list) and len(pt) == 2
                        for pt in segment)):
                raise ValueError(
                    "segment must have structure [[x1, y1], [x2, y2]]: " + name
                )
            (x1, y1), (x2, y2) = segment
            if not (x2 >= x1 and y2 >= y1):
                raise ValueError(
                    "segment must satisfy x2 >= x1 and y2 >= y1: " + name
                )
            if not (x1 == x2 or y1 == y2):
                raise ValueError(
                    "segment must be axis-aligned (horizontal or vertical): "
                    + name
                )
        if self.fluid:
            if "fluid" not in user_dict:
                raise ValueError("fluid missing in boundary: " + name)
            fluid_dict = user_dict["fluid"]
            if "type" not in fluid_dict:
                raise ValueError(
                    "type missing in fluid section of boundary: " + name
                )
            boundary_type = fluid_dict["type"]
            if boundary_type not in SUPPORTED_FLUID_BCS:
                raise ValueError(
                    "Unsupported boundary condition for fluid: " +
                    str(boundary_type)
                )
            fluid_config["type"] = boundary_type
            if boundary_type == "fixed_velocity":
                if "value" not in fluid_dict:
                    raise ValueError(
                        "value missing in fluid section of boundary: " + name
                    )
                value = fluid_dict["value"]
                if not isinstance(value, list) or len(value) != 2:
                    raise ValueError(
                        "value for fixed_velocity must be a list (ux, uy): "
                        + name
                    )
                fluid_config["vector_value"] = np.array(
                    value, dtype=control.precision)
            elif boundary_type == "fixed_pressure":
                if "value" not in fluid_dict:
                    raise ValueError(
                        "value missing in fluid section of boundary: " + name
                    )
                value = fluid_dict["value"]
                if not isinstance(value, (int, float)):
                    raise ValueError(
                        "value for fixed_pressure must be a float or int: "
                        + name
                    )
                fluid_config["scalar_value"] = control.precision(value)
        if fluid_config["type"] == "periodic":
            self.validate_periodic_boundary(name, segments, mesh)
        gx, gy = (int(v) for v in mesh.grid_global_shape)
        for segment_no, segment in enumerate(segments):
            (x1, y1), (x2, y2) = segment
            location = None
            if x2 - x1 == 0:
                orientation = "vertical"
                if x2 == gx - 1:
                    location = "right"
                elif x2 == 0:
                    location = "left"
            else:
                orientation = "horizontal"
                if y2 == gy - 1:
                    location = "top"
                elif y2 == 0:
                    location = "bottom"
            if location is None:
                # The reference fails here with an IndexError on its
                # location list (boundary.py:483-500).
                raise ValueError(
                    "segment does not lie on a domain side: " + name
                )
            self.boundary_elements.append(BoundaryElement(
                name, segment, segment_no, orientation, location, domain,
                control, fields, wall=wall, fluid_config=fluid_config))
        print_log("boundary name: " + name, domain.mpi_rank, verbose)

    def validate_periodic_boundary(self, name, segments, mesh):
        """boundary.py:521-603: two full-length segments on opposite
        sides."""
        if len(segments) != 2:
            raise ValueError(
                "For periodic boundary, each segment should be" +
                " followed by a corresponding periodic pair: " + name
            )
        s1, s2 = segments
        lx1 = abs(s1[1][0] - s1[0][0])
        ly1 = abs(s1[1][1] - s1[0][1])
        lx2 = abs(s2[1][0] - s2[0][0])
        ly2 = abs(s2[1][1] - s2[0][1])
        is_horizontal = ly1 == 0
        if is_horizontal and ly2 != 0:
            raise ValueError(
                "invalid periodic pair - different orientation: " + name
            )
        is_vertical = lx1 == 0
        if is_vertical and lx2 != 0:
            raise ValueError(
                "invalid periodic pair - different orientation: " + name
            )
        gx, gy = (int(v) for v in mesh.grid_global_shape)
        (x1_min, y1_min), (x1_max, y1_max) = s1
        (x2_min, y2_min), (x2_max, y2_max) = s2
        if is_horizontal:
            if ((x1_min, x1_max) != (0, gx - 1) or
                    (x2_min, x2_max) != (0, gx - 1)):
                raise ValueError(
                    "horizontal periodic boundaries must span" +
                    " the entire x-direction: " + name
                )
            if sorted([y1_min, y2_min]) != [0, gy - 1]:
                raise ValueError(
                    "horizontal periodic boundaries must connect" +
                    " top-bottom boundaries: " + name
                )
        elif is_vertical:
            if ((y1_min, y1_max) != (0, gy - 1) or
                    (y2_min, y2_max) != (0, gy - 1)):
                raise ValueError(
                    "vertical periodic boundaries must span" +
                    " the entire y-direction: " + name
                )
            if sorted([x1_min, x2_min]) != [0, gx - 1]:
                raise ValueError(
                    "vertical periodic boundaries must connect" +
                    " left-right boundaries: " + name
                )
        if is_vertical:
            self.x_periodic = True
        if is_horizontal:
            self.y_periodic = True
