"""Boundary elements on every side and in every combination the workloads
do not use, at sizes that cross the block widths (63, 65, 129 cells in y),
against the oracle.

The workloads put the moving wall on top, pressure ends and zero-gradient
walls on x / y respectively, and bounce-back on the walls.  Here each kind
also sits on the other sides: a moving wall on the left (x normal), a
channel periodic in y with bounce-back walls at x, pressure ends at y with
zero-gradient walls at x, a velocity inlet with a zero-gradient outlet, and
an obstacle that wraps across both periodic edges.  The initial velocity
varies from cell to cell (amplitude 0.03), so a wrong direction, sign or
neighbour shows far above the tolerances.

FP64: bar 1e-12 max relative error (expected bitwise).  FP32 / FP16 / BF16
storage: max |f - f_oracle| <= 2e-6 / 1e-4 / 1e-3 after 20 steps, velocity
within ten times that.
"""
import numpy as np
import pytest

from oracle import d2q9_oracle as O
from pylabolt_amd import cases
from tests.helpers import max_rel, oracle_case

pytestmark = pytest.mark.gpu


def _rough(sim):
    sim.initial_fields_dict["default"]["fluid"]["velocity"] = {
        "type": "func",
        "func": lambda i, j: (0.03 * np.sin(1.3 * i + 2.1 * j),
                              0.03 * np.cos(0.7 * i - 1.9 * j))}
    return sim


def _seg_x(i, ny):
    return [[i, 0], [i, ny - 1]]


def _seg_y(j, nx):
    return [[0, j], [nx - 1, j]]


def lid_left(nx=65, ny=70):
    sim = cases.lid_driven_cavity(nx, ny)
    sim.boundary_dict = {
        "options": {},
        "walls": {"wall": True,
                  "segments": [_seg_x(nx - 1, ny), _seg_y(0, nx),
                               _seg_y(ny - 1, nx)],
                  "fluid": {"type": "bounce_back"}},
        "lid": {"wall": True, "segments": [_seg_x(0, ny)],
                "fluid": {"type": "fixed_velocity", "value": [0.0, 0.08]}},
    }
    return sim


def channel_y(nx=33, ny=129):
    sim = cases.poiseuille_channel(nx, ny, model="BGK",
                                   forcing="guo_second_order")
    sim.boundary_dict = {
        "options": {},
        "ends": {"wall": False,
                 "segments": [_seg_y(0, nx), _seg_y(ny - 1, nx)],
                 "fluid": {"type": "periodic"}},
        "walls": {"wall": True,
                  "segments": [_seg_x(0, ny), _seg_x(nx - 1, ny)],
                  "fluid": {"type": "bounce_back"}},
    }
    sim.forcing_dict = {"gravity": [3e-6, 2e-5]}
    return sim


def pressure_y(nx=40, ny=65):
    sim = cases.cylinder(nx, ny, diameter=10, xc=20)
    sim.boundary_dict = {
        "options": {},
        "sides": {"wall": False,
                  "segments": [_seg_x(0, ny), _seg_x(nx - 1, ny)],
                  "fluid": {"type": "zero_gradient"}},
        "bottom": {"wall": False, "segments": [_seg_y(0, nx)],
                   "fluid": {"type": "fixed_pressure", "value": 1.0005}},
        "top": {"wall": False, "segments": [_seg_y(ny - 1, nx)],
                "fluid": {"type": "fixed_pressure", "value": 0.9995}},
    }
    sim.obstacle_dict["cylinder"]["center"] = [20.0, 30.0]
    return sim


def inlet_outlet(nx=50, ny=63):
    sim = cases.lid_driven_cavity(nx, ny, model="MRT")
    sim.boundary_dict = {
        "options": {},
        "walls": {"wall": True,
                  "segments": [_seg_y(0, nx), _seg_y(ny - 1, nx)],
                  "fluid": {"type": "bounce_back"}},
        "outlet": {"wall": False, "segments": [_seg_x(nx - 1, ny)],
                   "fluid": {"type": "zero_gradient"}},
        "inlet": {"wall": False, "segments": [_seg_x(0, ny)],
                  "fluid": {"type": "fixed_velocity", "value": [0.05, 0.0]}},
    }
    return sim


def periodic_wrap_obstacle(nx=48, ny=130):
    sim = cases.poiseuille_channel(nx, ny, model="BGK", forcing="guo_linear")
    sim.boundary_dict = {
        "options": {},
        "x": {"wall": False,
              "segments": [_seg_x(0, ny), _seg_x(nx - 1, ny)],
              "fluid": {"type": "periodic"}},
        "y": {"wall": False,
              "segments": [_seg_y(0, nx), _seg_y(ny - 1, nx)],
              "fluid": {"type": "periodic"}},
    }
    sim.obstacle_dict = {"options": {},
                         "post": {"type": "circle", "radius": 7.0,
                                  "center": [2.0, 3.0], "density": 1.0,
                                  "static": True}}
    sim.forcing_dict = {"gravity": [2e-5, -4e-6]}
    return sim


CASES = {f.__name__: f for f in (lid_left, channel_y, pressure_y,
                                 inlet_outlet, periodic_wrap_obstacle)}


def _solver(sim, storage, graph_steps=4):
    from pylabolt_amd.solvers.fluidLB import Solver
    s = Solver(simulation=sim, storage=storage, graph_steps=graph_steps,
               verbose=False)
    s.set_backend()
    return s


@pytest.mark.parametrize("name", sorted(CASES))
def test_boundaries_fp64_exact(name):
    sim = _rough(CASES[name]())
    case = oracle_case(sim)
    O.initialize_pop(case)
    s = _solver(sim, "f64")
    done = 0
    for chunk in (1, 4, 13):
        s.advance(chunk)
        O.run(case, chunk)
        done += chunk
        fl = case.fluid
        rho, vel = s.engine.moments_host()
        errs = {"pop": max_rel(s.engine.pop()[fl], case.pop[fl]),
                "pop_new": max_rel(s.engine.pop_new()[fl], case.pop_new[fl]),
                "density": max_rel(rho[fl], case.density[fl]),
                "velocity": max_rel(vel[fl], case.velocity[fl])}
        print(f"[boundaries] {name} step {done}: {errs}", flush=True)
        for k, v in errs.items():
            assert v <= 1e-12, f"{name} step {done}: {k} {v}"


# Max absolute error over every fluid cell and direction after 20 steps of
# a rough field (|u| up to 0.042, moving walls up to 0.1).  FP16: the
# largest deviations (w |3 c.u| up to ~0.03) sit on a grid of at most
# 2^-16 ~ 1.5e-5; 20 independent stochastic roundings of up to one grid
# step each, with the tail of the max over ~10^4 values, stay within
# 2e-4.  BF16 has 3 bits fewer.
ABS_TOL = {"f32": 2e-6, "f16dev": 2e-4, "bf16dev": 1e-3}


@pytest.mark.parametrize("storage", sorted(ABS_TOL))
@pytest.mark.parametrize("name", sorted(CASES))
def test_boundaries_reduced_precision(name, storage):
    sim = _rough(CASES[name]())
    case = oracle_case(sim)
    O.initialize_pop(case)
    s = _solver(sim, storage)
    s.advance(20)
    O.run(case, 20)
    fl = case.fluid
    err_pop = float(np.abs(s.engine.pop()[fl] - case.pop[fl]).max())
    err_new = float(np.abs(s.engine.pop_new()[fl] - case.pop_new[fl]).max())
    _, vel = s.engine.moments_host()
    err_u = float(np.abs(vel[fl] - case.velocity[fl]).max())
    print(f"[boundaries] {name} {storage}: pop {err_pop:.2e} "
          f"pop_new {err_new:.2e} u {err_u:.2e}", flush=True)
    tol = ABS_TOL[storage]
    assert err_pop <= tol and err_new <= tol and err_u <= 10 * tol
