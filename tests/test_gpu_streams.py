"""Two host threads, two lattices, two streams.

The sweeps fork part of each step onto an auxiliary stream and join it
back; the library keeps one auxiliary stream and event pair per caller
stream, so lattices driven from different threads on different streams do
not order each other's work.  Each lattice advanced concurrently must equal,
bit for bit, the same lattice advanced alone.
"""
import threading

import numpy as np
import pytest

from pylabolt_amd import cases

pytestmark = pytest.mark.gpu


def _solver(sim, two_step):
    from pylabolt_amd.solvers.fluidLB import Solver
    s = Solver(simulation=sim, verbose=False, graph_steps=4,
               two_step=two_step)
    s.set_backend()
    return s


def _cases():
    cyl = cases.cylinder(44, 130, diameter=10, xc=14)   # bulk + general tiles
    chan = cases.poiseuille_channel(120, 300, model="BGK",
                                    forcing="guo_second_order")
    chan.forcing_dict["gravity"] = [2e-5, 0.0]
    return [(cyl, False), (chan, True)]


def _state(solver):
    eng = solver.engine
    eng.synchronize()
    rho, vel = eng.moments_host()
    return eng.pop(), rho, vel


def _drive(solver, rounds, barrier=None):
    if barrier is not None:
        barrier.wait()
    for r in range(rounds):
        # eager steps (fork/join per call) and graph replays, interleaved
        solver.engine.advance(2, use_graph=False)
        solver.engine.advance(9)


def test_two_threads_on_two_streams_bitwise():
    rounds = 12
    alone = []
    for sim, two in _cases():
        s = _solver(sim, two)
        _drive(s, rounds)
        alone.append(_state(s))
    solvers = [_solver(sim, two) for sim, two in _cases()]
    assert solvers[1].engine.two_step, "the channel must use the two-step sweep"
    assert solvers[0].engine.stream.cuda_stream != \
        solvers[1].engine.stream.cuda_stream
    barrier = threading.Barrier(2)
    errors = []

    def run(s):
        try:
            _drive(s, rounds, barrier)
        except Exception as e:  # reported below
            errors.append(e)

    threads = [threading.Thread(target=run, args=(s,)) for s in solvers]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert not errors, errors
    for n, s in enumerate(solvers):
        got = _state(s)
        for a, b, what in zip(got, alone[n], ("pop", "density", "velocity")):
            assert np.array_equal(a, b), f"lattice {n}: {what} differs"
    print("[streams] two threads, two streams: bitwise equal to alone",
          flush=True)
