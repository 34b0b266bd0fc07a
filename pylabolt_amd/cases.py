"""Case files for the standard workloads, as ``simulation`` namespaces.

Each function returns what a user's ``simulation.py`` would define (the
reference's dict API, snake_case, see pylabolt/solvers/fluidLB.py and the
``*_dict`` readers in pylabolt/base/).  The sizes default to the full
workloads; tests pass smaller grids.

* ``lid_driven_cavity``: D2Q9 BGK, walls bounce_back, lid fixed_velocity
  listed last so it wins at the two top corners.
* ``poiseuille_channel``: periodic in x, bounce_back walls in y, uniform
  body force ``g_x = 8 nu u_max / H**2`` with Guo forcing.
* ``cylinder``: static circle of diameter ``D`` at ``(x_c, y_c + delta)``,
  zero_gradient top/bottom listed first, then fixed_pressure inlet/outlet
  (so the pressure elements win at the four corners), ``nu`` for ``Re``
  at reference speed ``u_ref``.
"""
from types import SimpleNamespace

import numpy as np


def _control(steps, precision="double", std_out=None, save=None):
    return {"start_time": 0, "end_time": int(steps),
            "std_out_interval": std_out, "save_interval": save,
            "checkpoint_interval": None, "precision": precision}


def _fluid_init(rho=1.0, u=(0.0, 0.0)):
    return {"default": {"fluid": {
        "velocity": {"type": "fixed", "value": [float(u[0]), float(u[1])]},
        "density": {"type": "fixed", "value": float(rho)},
        "pressure": {"type": "fixed", "value": float(rho) / 3.0},
    }}}


def lid_driven_cavity(nx=128, ny=128, re=100.0, u_lid=0.1, steps=10000,
                      model="BGK", precision="double"):
    """Lid-driven cavity, tau = 0.5 + 3 u (N - 1) / Re."""
    kin_visc = u_lid * (nx - 1) / re
    return SimpleNamespace(
        control_dict=_control(steps, precision),
        mesh_dict={"grid": [nx, ny]},
        lattice_dict={"lattice_type": "D2Q9"},
        decompose_dict={"nx": 1, "ny": 1},
        transport_dict={"kin_visc": kin_visc},
        initial_fields_dict=_fluid_init(),
        boundary_dict={
            "options": {},
            "walls": {"wall": True,
                      "segments": [[[0, 0], [0, ny - 1]],
                                   [[nx - 1, 0], [nx - 1, ny - 1]],
                                   [[0, 0], [nx - 1, 0]]],
                      "fluid": {"type": "bounce_back"}},
            "lid": {"wall": True,
                    "segments": [[[0, ny - 1], [nx - 1, ny - 1]]],
                    "fluid": {"type": "fixed_velocity",
                              "value": [u_lid, 0.0]}},
        },
        obstacle_dict={"options": {}},
        collision_dict={"fluid": {"model": model,
                                  "equilibrium": "density_based_second_order",
                                  "forcing_model": None}},
        forcing_dict={},
    )


def poiseuille_channel(nx=4096, ny=1024, re=100.0, u_max=0.05, steps=20000,
                       model="MRT", forcing="guo_second_order",
                       precision="double"):
    """Gravity-driven channel: nu = u_max H / Re, g = 8 nu u_max / H^2."""
    h = float(ny)
    kin_visc = u_max * h / re
    gx = 8.0 * kin_visc * u_max / (h * h)
    return SimpleNamespace(
        control_dict=_control(steps, precision),
        mesh_dict={"grid": [nx, ny]},
        lattice_dict={"lattice_type": "D2Q9"},
        decompose_dict={"nx": 1, "ny": 1},
        transport_dict={"kin_visc": kin_visc},
        initial_fields_dict=_fluid_init(),
        boundary_dict={
            "options": {},
            "inlet_outlet": {"wall": False,
                             "segments": [[[0, 0], [0, ny - 1]],
                                          [[nx - 1, 0], [nx - 1, ny - 1]]],
                             "fluid": {"type": "periodic"}},
            "walls": {"wall": True,
                      "segments": [[[0, 0], [nx - 1, 0]],
                                   [[0, ny - 1], [nx - 1, ny - 1]]],
                      "fluid": {"type": "bounce_back"}},
        },
        obstacle_dict={"options": {}},
        collision_dict={"fluid": {"model": model,
                                  "equilibrium": "density_based_second_order",
                                  "forcing_model": forcing}},
        forcing_dict={"gravity": [gx, 0.0]},
    )


def cylinder_offset(seed=0):
    """Seeded vertical offset of the cylinder centre, in [-4, 4]."""
    return int(np.random.default_rng(seed).integers(-4, 5))


def cylinder(nx=8192, ny=2048, diameter=128, xc=None, re=100.0, u_ref=0.05,
             delta_rho=1e-3, seed=0, steps=50000, model="BGK",
             precision="double"):
    """Flow past a static cylinder between fixed-pressure ends."""
    xc = nx // 4 if xc is None else xc
    yc = ny // 2 + cylinder_offset(seed)
    kin_visc = u_ref * diameter / re
    return SimpleNamespace(
        control_dict=_control(steps, precision),
        mesh_dict={"grid": [nx, ny]},
        lattice_dict={"lattice_type": "D2Q9"},
        decompose_dict={"nx": 1, "ny": 1},
        transport_dict={"kin_visc": kin_visc},
        initial_fields_dict=_fluid_init(),
        boundary_dict={
            "options": {},
            "top_bottom": {"wall": False,
                           "segments": [[[0, 0], [nx - 1, 0]],
                                        [[0, ny - 1], [nx - 1, ny - 1]]],
                           "fluid": {"type": "zero_gradient"}},
            "inlet": {"wall": False, "segments": [[[0, 0], [0, ny - 1]]],
                      "fluid": {"type": "fixed_pressure",
                                "value": 1.0 + 0.5 * delta_rho}},
            "outlet": {"wall": False,
                       "segments": [[[nx - 1, 0], [nx - 1, ny - 1]]],
                       "fluid": {"type": "fixed_pressure",
                                 "value": 1.0 - 0.5 * delta_rho}},
        },
        obstacle_dict={"options": {},
                       "cylinder": {"type": "circle",
                                    "radius": diameter / 2.0,
                                    "center": [float(xc), float(yc)],
                                    "density": 1.0, "static": True}},
        collision_dict={"fluid": {"model": model,
                                  "equilibrium": "density_based_second_order",
                                  "forcing_model": None}},
        forcing_dict={},
    )


def adjoint_channel(nx=1024, ny=256, steps=2000, kin_visc=0.1,
                    forcing="guo_second_order", seed=0):
    """Guo-forced channel for the adjoint: x periodic, bounce-back walls at
    y, per-cell body force F_x ~ N(1e-6, 2e-7) (F_y = 0) from ``seed``,
    initial velocity zero, and a seeded target parabola for u_x on the
    outlet column.  Returns (simulation, force (nx, ny, 2), u0 (nx, ny, 2),
    target (ny,))."""
    sim = poiseuille_channel(nx, ny, model="BGK", forcing=forcing,
                             steps=steps)
    sim.transport_dict["kin_visc"] = float(kin_visc)
    sim.forcing_dict = {}
    rng = np.random.default_rng(seed)
    force = np.zeros((nx, ny, 2))
    force[:, :, 0] = rng.normal(1e-6, 2e-7, size=(nx, ny))
    u0 = np.zeros((nx, ny, 2))
    y = (np.arange(ny) + 0.5) / ny
    amp = 1e-3 * (1.0 + 0.1 * rng.standard_normal())
    target = amp * 4.0 * y * (1.0 - y)
    return sim, force, u0, target


def batch_cases(n=256, seed=0, nx_range=(64, 1024), ny_range=(64, 512),
                step=32, re_range=(50.0, 400.0), steps=5000,
                precision="single"):
    """``n`` independent lattices for a batched sweep, drawn from ``seed``:
    sizes uniform over the multiples of ``step`` in the ranges, Re uniform
    in ``re_range``; even ids are lid-driven cavities (BGK), odd ids
    gravity-driven channels (MRT, second-order Guo forcing)."""
    rng = np.random.default_rng(seed)
    xs = np.arange(nx_range[0], nx_range[1] + 1, step)
    ys = np.arange(ny_range[0], ny_range[1] + 1, step)
    out = []
    for k in range(n):
        nx = int(rng.choice(xs))
        ny = int(rng.choice(ys))
        re = float(rng.uniform(*re_range))
        if k % 2 == 0:
            sim = lid_driven_cavity(nx, ny, re=re, steps=steps, model="BGK",
                                    precision=precision)
        else:
            sim = poiseuille_channel(nx, ny, re=re, steps=steps, model="MRT",
                                     forcing="guo_second_order",
                                     precision=precision)
        out.append(sim)
    return out
