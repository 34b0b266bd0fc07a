"""Write the golden fixtures: the oracle's state after a few steps of tiny
cases that cover every boundary type, the corners where two elements meet,
periodic wraps, an obstacle, both collision models and all forcings.

Run from the repository root:  python tests/golden/make_golden.py
Each case is one .npz: the padded-grid populations (pop, pop_new), density
and velocity after ``steps`` time steps, plus the fluid mask.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import d2q9_oracle as O  # noqa: E402
from pylabolt_amd import cases  # noqa: E402
from tests.helpers import oracle_case  # noqa: E402


def golden_cases():
    out = {}
    out["cavity_bgk_7x6"] = (cases.lid_driven_cavity(7, 6), 30)
    out["cavity_mrt_6x7"] = (cases.lid_driven_cavity(6, 7, model="MRT"), 30)
    ch = cases.poiseuille_channel(8, 5, model="BGK", forcing="guo_second_order")
    ch.forcing_dict["gravity"] = [3e-5, 1e-5]
    out["channel_bgk_guo2_8x5"] = (ch, 30)
    ch = cases.poiseuille_channel(7, 6, model="MRT", forcing="guo_linear")
    ch.forcing_dict["gravity"] = [2e-5, -1e-5]
    out["channel_mrt_guo1_7x6"] = (ch, 30)
    out["cylinder_zg_pres_24x14"] = (
        cases.cylinder(24, 14, diameter=6, xc=8, delta_rho=0.02), 30)
    box = cases.poiseuille_channel(6, 5, model="BGK", forcing="guo_linear")
    box.boundary_dict = {
        "options": {},
        "x": {"wall": False, "segments": [[[0, 0], [0, 4]], [[5, 0], [5, 4]]],
              "fluid": {"type": "periodic"}},
        "y": {"wall": False, "segments": [[[0, 0], [5, 0]], [[0, 4], [5, 4]]],
              "fluid": {"type": "periodic"}},
    }
    box.forcing_dict["gravity"] = [1e-4, 2e-4]
    box.initial_fields_dict["default"]["fluid"]["velocity"] = {
        "type": "func", "func": lambda i, j: (0.01 * (i - 2), 0.02 * (j - 2))}
    out["periodic_box_6x5"] = (box, 20)
    return out


def main():
    for name, (sim, steps) in golden_cases().items():
        case = oracle_case(sim)
        O.initialize_pop(case)
        O.run(case, steps)
        np.savez_compressed(os.path.join(HERE, name + ".npz"),
                            steps=steps, pop=case.pop, pop_new=case.pop_new,
                            density=case.density, velocity=case.velocity,
                            fluid=case.fluid)
        print(name, "steps", steps)


if __name__ == "__main__":
    main()
