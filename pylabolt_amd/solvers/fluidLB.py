"""fluidLB solver on one MI355X (reference: pylabolt/solvers/fluidLB.py).

``FluidLB`` carries the reference's model tables (fluidLB.py:22-60);
``Solver`` keeps its interface -- ``set_backend``, ``compile``,
``single_time_step``, ``run``, ``verify_kernel_signatures`` and the
operator attributes -- with the time step executed by the fused HIP sweep:

* reference step (fluidLB.py:206-253): density, force, velocity, collide,
  periodic halo, stream, boundaries -- seven or more full-grid passes;
* here: one sweep per step that streams (with the periodic wrap and every
  boundary rule) and then computes moments and collides, replayed from a
  HIP graph in chunks (the reference captures one step in a CUDA graph,
  fluidLB.py:270-287).

After any number of steps the solver presents exactly the reference's
state: ``pop_fluid`` = post-collision populations of the last step,
``pop_fluid_new`` = populations after its stream and boundaries,
``density``/``velocity`` = the moments that step computed
(``sync_fields()`` copies them into ``state.fields``).
"""
import time

import numpy as np

from pylabolt_amd import _native as nt
from pylabolt_amd.backend import Backend
from pylabolt_amd.base.collision_operator import CollisionOperator
from pylabolt_amd.base.force_operator import ForceOperator
from pylabolt_amd.base.operators import (BoundaryOperator,
                                         ComputeFieldsOperator, MPIOperator,
                                         ObstacleOperator, StreamingOperator)
from pylabolt_amd.base.state import State
from pylabolt_amd.utils.helpers import (SimulationStatusLogger, SingleRankComm,
                                        load_simulation, print_log)
from pylabolt_amd.utils.io_operator import InputOutputOperator
from pylabolt_amd.utils.residues import ResidueOperator


class FluidLB:
    """Model tables of the reference solver (fluidLB.py:22-60)."""

    def __init__(self):
        self.solver_name = "fluidLB"
        self.equilibrium_models = {"fluid": ["density_based_second_order"]}
        self.forcing_models = {"fluid": [None, "guo_linear",
                                         "guo_second_order"]}
        self.collision_models = {"fluid": ["BGK", "MRT"]}
        self.streaming_type = {"fluid": "scalar_based"}
        self.boundary_condition_type = {"fluid": "density_based"}
        self.compute_fields_config = {"type": "density_based",
                                      "moment_fields": ["density", "velocity"]}
        self.obstacle_kernels_type = "single_phase"
        self.residue_fields = ["density", "velocity"]
        self.save_fields = ["density", "velocity", "solid", "solid_id",
                            "solid_boundary", "fluid_boundary",
                            "surface_normals"]


class Solver:
    """Drop-in for the reference ``Solver`` (fluidLB.py:107-395).

    ``simulation`` defaults to ``simulation.py`` in the working directory,
    as in the reference.  Extra keywords select the population storage
    (``"f64"`` is the reference-exact path), the launch tile, the HIP
    graph chunk, the torch stream the engine runs on (a new one if None)
    how tiles holding solid cells are swept (``solid_tiles``: "sweep",
    "cells" or "auto"; None picks per storage form) and whether pairs of
    steps run as one sweep over the deep interior (``two_step``; None: on
    for FP64 and FP32 storage when the two population buffers exceed 64
    MiB, off for 16-bit storage).  Every choice gives bitwise the same
    results, but for ``two_step`` in 16-bit storage, whose intermediate
    step stays in FP32.
    """

    def __init__(self, comm=None, backend="gpu", n_threads=None,
                 simulation=None, storage="f64", tile=(64, 4),
                 graph_steps=16, verbose=True, output_root=".", stream=None,
                 solid_tiles=None, two_step=None):
        comm = comm if comm is not None else SingleRankComm()
        mpi_rank = comm.Get_rank()
        self.verbose = verbose
        print_log(f"{'Solver':<10}: fluidLB (MI355X)", mpi_rank, verbose)
        self.model = FluidLB()
        if simulation is None:
            simulation = load_simulation(comm, mpi_rank)
        self.simulation = simulation
        self.state = State(simulation, comm, mpi_rank, fluid=True)
        self.backend = Backend(comm, self.state, backend, n_threads,
                               storage=storage, tile=tile,
                               graph_steps=graph_steps, stream=stream,
                               solid_tiles=solid_tiles, two_step=two_step)
        self.mpi_operator = MPIOperator(comm, self.state)
        self.collision_operator = CollisionOperator(simulation, self.model,
                                                    self.state, comm,
                                                    verbose=verbose)
        self.force_operator = ForceOperator(
            simulation, self.model, self.state, comm,
            collision_operator=self.collision_operator, verbose=verbose)
        self.compute_fields_operator = ComputeFieldsOperator(self.model,
                                                             self.state)
        self.streaming_operator = StreamingOperator(self.model, self.state,
                                                    self.backend)
        self.boundary_operator = BoundaryOperator(self.model, self.state,
                                                  self.backend)
        self.obstacle_operator = ObstacleOperator(self.model, self.state,
                                                  self.backend,
                                                  self.mpi_operator,
                                                  self.force_operator)
        self.residue_operator = ResidueOperator(self.model, self.state,
                                                self.backend, comm)
        self.io_operator = InputOutputOperator(self.model, self.state,
                                               self.backend, comm,
                                               root=output_root)
        self.logger = SimulationStatusLogger(mpi_rank, verbose=verbose)
        self.execute_single_time_step = self.single_time_step
        self.run_time = None

    # ---- reference interface ---------------------------------------------
    def set_backend(self, verbose=None):
        """Build the device state (fluidLB.py:186-204)."""
        self.backend.create_engine(self.state, self.collision_operator,
                                   self.force_operator)
        eng = self.backend.engine
        print_log(f"{'Engine':<10}: storage {eng.storage_name}, two-step "
                  f"sweep {'on' if eng.two_step else 'off'}",
                  0, self.verbose if verbose is None else verbose)
        for op in (self.mpi_operator, self.collision_operator,
                   self.force_operator, self.compute_fields_operator,
                   self.streaming_operator, self.boundary_operator,
                   self.obstacle_operator, self.residue_operator,
                   self.io_operator):
            op.set_backend(self.state, self.backend)

    @property
    def engine(self):
        if self.backend.engine is None:
            raise RuntimeError("call set_backend() first")
        return self.backend.engine

    def compile(self, verbose=None):
        """The reference JIT-compiles each kernel and captures a CUDA graph
        of one step (fluidLB.py:255-287).  The kernels here are compiled
        ahead of time; graphs are captured on first use by ``advance``."""
        self.execute_single_time_step = self.single_time_step

    def single_time_step(self, verbose=True):
        """One reference time step (fluidLB.py:206-253), obstacle forces
        included when they are asked for."""
        self.engine.step(write_moments=True)
        self.obstacle_operator.compute_force_torque(self.state, self.backend,
                                                    self.mpi_operator)

    def advance(self, steps, write_moments=True):
        """``steps`` time steps, replaying HIP graphs between the ends."""
        self.engine.advance(steps, write_moments=write_moments)

    def run(self, verbose=True):
        """The reference run loop (fluidLB.py:309-395): initialise, then
        step until ``end_time``, with residues, status lines and field
        output at their intervals.  Steps between two such events run as
        one ``advance`` call."""
        st = self.state
        self.collision_operator.initialize_pop(st, self.backend)
        self.residue_operator.set_backend(st, self.backend)
        self.boundary_operator.compute_force(st, self.backend,
                                             self.mpi_operator, time_step=0)
        self.obstacle_operator.compute_force_torque(st, self.backend,
                                                    self.mpi_operator)
        self.io_operator.write_fields(st, self.backend, st.control.start_time,
                                      solver=self)
        self.io_operator.write_histories(st, time_step=0)
        t0 = time.perf_counter()
        t = st.control.start_time
        end = st.control.end_time
        while t < end:
            nxt = self._next_event(t, end)
            self.advance(nxt - t, write_moments=True)
            t = nxt
            self.boundary_operator.compute_force(st, self.backend,
                                                 self.mpi_operator, t)
            if self._obstacle_event(t):
                self.obstacle_operator.compute_force_torque(
                    st, self.backend, self.mpi_operator)
            self.residue_operator.compute_residues(st, self.backend,
                                                   self.mpi_operator, t)
            self.logger.log_data(
                st, t, res_density=self.residue_operator.residues["res_density"],
                res_velocity=self.residue_operator.residues["res_velocity"])
            self.io_operator.write_fields(st, self.backend, t, solver=self)
            self.io_operator.write_histories(st, t)
        if end > st.control.start_time and not self._obstacle_event(end):
            # the reference leaves the last step's obstacle forces in place
            # (it evaluates them every step); the loop above evaluated them
            # at the write steps only
            self.obstacle_operator.compute_force_torque(st, self.backend,
                                                        self.mpi_operator)
        self.engine.synchronize()
        self.run_time = time.perf_counter() - t0
        print_log(f"{'Simulation complete, run time':<30}: "
                  f"{str(self.run_time) + ' s':<30}", 0, self.verbose)

    def _event_intervals(self):
        st = self.state
        ivs = [st.control.std_out_interval, st.control.save_interval]
        if st.boundary.compute_force:
            ivs.append(st.boundary.write_interval)
        if st.obstacle.compute_force_torque and st.obstacle.write_obstacle_data:
            ivs.append(st.obstacle.write_interval)
        return [iv for iv in ivs if iv]

    def _obstacle_event(self, t):
        """Obstacle forces only feed the histories for static bodies: they
        are evaluated at the obstacle write steps (the reference evaluates
        them every step, fluidLB.py:249-253)."""
        obs = self.state.obstacle
        return (obs.compute_force_torque and obs.write_obstacle_data and
                obs.write_interval and t % obs.write_interval == 0)

    def _next_event(self, t, end):
        nxt = end
        for iv in self._event_intervals():
            nxt = min(nxt, (t // iv + 1) * iv)
        return nxt

    def verify_kernel_signatures(self, verbose=True):
        """The reference detects Numba recompiles (fluidLB.py:289-307).  The
        kernels here are compiled ahead of time; this checks that the
        library still exports every entry point the engine binds."""
        lib = nt.lib()
        for name in nt.SIGNATURES:
            if not hasattr(lib, name):
                raise RuntimeError(f"Developer error! {name} missing")

    # ---- host views --------------------------------------------------------
    def sync_fields(self):
        """Copy the device state into ``state.fields`` in the reference's
        layouts (pop_fluid, pop_fluid_new, density, velocity, force_field)."""
        eng = self.engine
        f = self.state.fields
        f.pop_fluid[:] = eng.pop()
        f.pop_fluid_new[:] = eng.pop_new()
        if eng.mom_full:
            rho, vel = eng.moments_host()
            fluid = ~f.solid & ~f.ghost_node
            f.density[fluid] = rho[fluid]
            f.velocity[fluid] = vel[fluid]
            f.force_field[:] = self.force_operator.force_field_host(self.state)
        return f


def main(backend="gpu", n_threads=None, debug_mode=False, **kw):
    """Reference ``fluidLB.main`` (fluidLB.py:398-407) without MPI."""
    solver = Solver(None, backend, n_threads, **kw)
    solver.set_backend()
    solver.compile()
    solver.run()
    if debug_mode:
        solver.verify_kernel_signatures()
    return solver
