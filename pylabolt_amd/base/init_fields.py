"""Initial fields (reference: pylabolt/base/init_fields.py:7-375).

``initial_fields_dict = {default: {fluid: {velocity, density, pressure}},
<region>: {fluid: {...}}}``; each entry is ``{type: fixed, value}`` or
``{type: func, func}``.  Values are written on non-ghost cells only; a
``func`` receives global indices ``(i_global, j_global)``.  Region entries
override in dict order (they cover the whole sub-domain, as in the
reference).
"""
import numpy as np

from pylabolt_amd.base.domain import local_to_global
from pylabolt_amd.utils.helpers import print_log


def init_fields(simulation, control, domain, fields, fluid=True, verbose=True):
    print_log("-" * 80, domain.mpi_rank, verbose)
    print_log("Initializing fields...\n", domain.mpi_rank, verbose)
    if not hasattr(simulation, "initial_fields_dict"):
        raise ValueError("initial_fields_dict not found in simulation.py file")
    initial_fields_dict = simulation.initial_fields_dict
    if "default" not in initial_fields_dict:
        raise ValueError("default missing in initial_fields_dict")
    default_dict = initial_fields_dict["default"]
    if fluid:
        if "fluid" not in default_dict:
            raise ValueError("fluid missing in initial_fields_dict - default")
        init_fields_fluid(default_dict["fluid"], domain, fields, control,
                          default=True)
    for key in initial_fields_dict.keys():
        if key == "default":
            continue
        user_dict = initial_fields_dict[key]
        if "fluid" in user_dict and fluid:
            init_fields_fluid(user_dict["fluid"], domain, fields, control,
                              default=False)
    print_log("Initializing fields done!", domain.mpi_rank, verbose)
    print_log("-" * 80, domain.mpi_rank, verbose)


def init_fields_fluid(fluid_dict, domain, fields, control, default=True):
    if default:
        try:
            velocity_dict = fluid_dict["velocity"]
            density_dict = fluid_dict["density"]
            pressure_dict = fluid_dict["pressure"]
        except KeyError as e:
            raise ValueError("" + str(e) + " is missing in default")
        read_dict(velocity_dict, fields.velocity, fields.ghost_node, domain,
                  control, scalar_dict=False)
        read_dict(density_dict, fields.density, fields.ghost_node, domain,
                  control, scalar_dict=True)
        read_dict(pressure_dict, fields.pressure, fields.ghost_node, domain,
                  control, scalar_dict=True)
    else:
        for name, scalar in (("velocity", False), ("density", True),
                             ("pressure", True)):
            if name in fluid_dict:
                read_dict(fluid_dict[name], getattr(fields, name),
                          fields.ghost_node, domain, control,
                          scalar_dict=scalar)


def read_dict(dict_input, field, ghost_node, domain, control,
              scalar_dict=True):
    """Parse one field entry (init_fields.py:274-322), including the
    reference's operator precedence in its value check."""
    value = None
    func = None
    if "type" not in dict_input:
        raise ValueError("type missing in field definition")
    dict_type = dict_input["type"]
    if dict_type == "fixed":
        if "value" not in dict_input:
            raise ValueError("value missing for fixed type field definition")
        value = dict_input["value"]
        if not scalar_dict and type(value) is list and len(value) == 2:
            value = np.array(value, dtype=control.precision)
        elif scalar_dict and type(value) is float or type(value) is int:
            value = control.precision(value)
        else:
            raise ValueError(
                "vector value must be a list (ux, uy)" +
                " and scalar value must be a float or int"
            )
    elif dict_type == "func":
        if "func" not in dict_input:
            raise ValueError("func missing for func type field definition")
        func = dict_input["func"]
    else:
        raise ValueError("Unsupported velocity initialization")
    set_field(domain, field, ghost_node, value=value, func=func,
              vector=not scalar_dict)


def set_field(domain, field, ghost_node, value=None, func=None, vector=False):
    """Write ``value`` or ``func(i_global, j_global)`` on non-ghost cells
    (init_fields.py:325-375)."""
    if value is not None:
        if vector:
            field[~ghost_node, :] = value
        else:
            field[~ghost_node] = value
    elif func is not None:
        ny_pad = int(domain.shape[1])
        for ind in np.flatnonzero(~ghost_node):
            i = ind // ny_pad
            j = ind - i * ny_pad
            ig, jg = local_to_global(i - 1, j - 1, domain.offset)
            if vector:
                field[ind, 0], field[ind, 1] = func(ig, jg)
            else:
                field[ind] = func(ig, jg)
