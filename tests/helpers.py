"""Shared test helpers: build matching engine and oracle runs of a case."""
import numpy as np

from oracle import d2q9_oracle as O
from pylabolt_amd.base.collision_operator import CollisionOperator
from pylabolt_amd.base.force_operator import ForceOperator
from pylabolt_amd.base.state import State
from pylabolt_amd.solvers.fluidLB import FluidLB


def oracle_case(sim, force_field=None):
    """OracleCase of a simulation namespace (the package builds the case
    arrays; the oracle runs the steps)."""
    st = State(sim)
    co = CollisionOperator(sim, FluidLB(), st, verbose=False)
    fo = ForceOperator(sim, FluidLB(), st, collision_operator=co,
                       verbose=False)
    return O.case_from_state(
        st, co.collision_fluid, co.forcing_fluid, co.omega_fluid, fo.gravity,
        force_source="gravity" if force_field is None else "field",
        force_field=force_field)


def max_rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = np.maximum(np.abs(b), 1e-300)
    d = np.abs(a - b)
    nz = np.abs(b) > 0
    rel = np.where(nz, d / scale, d)
    return float(rel.max()) if rel.size else 0.0
