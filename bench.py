"""Benchmark: MLUPS and HBM roofline of the fused D2Q9 sweep on one MI355X.

Workload (default): flow past a cylinder, D2Q9 BGK, 8192 x 2048 interior
cells, cylinder D=128, fixed-pressure inlet/outlet, zero-gradient top and
bottom, FP64 populations -- the configuration the metric is quoted on.
Synthetic initial state (rest, rho = 1), as in the reference's case files.

One "step" is one reference time step (stream + boundaries + moments +
collision) over the whole lattice: one fused kernel launch, replayed from a
HIP graph.  MLUPS = Nx * Ny * steps / t / 1e6.  The kernel time comes from
HIP events recorded on the engine's own stream around the timed steps.

Roofline: a single-step sweep reads and writes the 9 populations of every
cell once per update: 144 B in FP64, 72 B in FP32, 36 B for 16-bit
storage.  FP64 and FP32 advance pairs of steps with one launch
(d2q9_fused_step2): the deep interior is read and written once per two
updates (144 B per cell and launch in FP64), the rim twice (288 B);
``achieved`` is those bytes over the launch's time (two steps),
``single_step_equivalent_gbs`` what a single-step sweep would have to move
at the measured update rate.  Peak is
MI355X's 8 TB/s HBM3E.  ``traffic`` is the measured HBM bytes per launch
from rocprofv3 FETCH_SIZE/WRITE_SIZE counters (profiles/), when present.

The CPU baseline is the C++/OpenMP restatement of the reference's
multi-pass step (oracle/d2q9_ref.cpp) on the same workload for a bounded
number of steps, on the host cores.

The timed steps replay HIP graphs only: the graph chunk is chosen to
divide ``--steps`` (a multiple of 4 up to 64; otherwise chunks of 32 and
the remainder eager) and captured before the timer starts.

Also reported: ``forms``, FP32 and 16-bit-storage runs of the same
workload with their roofline fractions and, after ``--long-steps`` steps
of each (50,000, the case's length, by default), the velocity error
against the FP64 run (relative L2 over the fluid cells, and RMS error over
the case's reference speed u_ref = 0.05); ``channel``, the gravity-driven
MRT + Guo channel 4096 x 1024 at its full 20,000 steps in every storage
form: MLUPS and roofline of each, velocity error against the FP64 run and
of every form against the start-up series of plane Poiseuille flow;
``batched``, 256 seeded lattices of uneven size advanced by one persistent
launch per step; ``adjoint``, the forward + backward cost of the adjoint
channel over its plain forward (one graph of all its steps), run in a
child process (``ChildLeg``) so that a failure there cannot take the
benchmark line with it; and a plain copy kernel's bandwidth on the same
device.
"""
import argparse
import importlib
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

BYTES_PER_CELL = {"f64": 144, "f32": 72, "f16dev": 36, "bf16dev": 36}
DTYPES = {"f64": "f64", "f32": "f32", "f16dev": "f32-compute/f16-store",
          "bf16dev": "f32-compute/bf16-store"}
PEAK_GBS = 8000.0
# reference speed of the cylinder case (Re = u_ref D / nu)
U_REF = 0.05


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def make_workload(name, nx, ny, steps):
    from pylabolt_amd import cases
    if name == "cylinder":
        return cases.cylinder(nx or 8192, ny or 2048, steps=steps)
    if name == "poiseuille":
        return cases.poiseuille_channel(nx or 4096, ny or 1024, steps=steps)
    if name == "cavity":
        return cases.lid_driven_cavity(nx or 128, ny or 128, steps=steps)
    raise ValueError(name)


def model_name(sim):
    """Collision model and forcing of a case dict, e.g. "MRT+guo_second_order"."""
    fl = sim.collision_dict["fluid"]
    forcing = fl.get("forcing_model")
    return fl["model"] + (f"+{forcing}" if forcing else "")


def graph_steps_for(steps):
    """Graph chunk for ``steps`` timed steps: the largest multiple of 4 up
    to 64 that divides it (every timed step then replays a graph), else
    32."""
    for g in range(64, 3, -4):
        if steps % g == 0:
            return g
    return 32


def busy_ahead(stream, nbytes=512 << 20):
    """Queue ~170 us of copy-kernel work on ``stream`` ahead of a timed
    region's start event, so that the host has submitted the region's graph
    by the time the event fires: the region then times the GPU's work, not
    the host's submission of the graph's nodes (about 10 us per step of a
    20-step region)."""
    import ctypes
    import torch
    from pylabolt_amd import _native as nt
    global _BUSY
    if _BUSY is None or _BUSY[0].device != torch.device("cuda",
                                                        torch.cuda.current_device()):
        a = torch.zeros(nbytes // 8, dtype=torch.float64, device="cuda")
        _BUSY = (a, torch.empty_like(a))
    a, b = _BUSY
    nt.check(nt.lib().d2q9_copy(nt.ptr(b), nt.ptr(a), nbytes,
                                ctypes.c_void_p(stream.cuda_stream)), "copy")


_BUSY = None


def time_engine(solver, steps, warmup, use_graph=True):
    import torch
    eng = solver.engine
    eng.advance(warmup, write_moments=False, use_graph=use_graph)
    if use_graph:
        eng.capture_graph()
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    if use_graph:
        busy_ahead(eng.stream)
    t0 = time.perf_counter()
    start.record(eng.stream)
    eng.advance(steps, write_moments=False, use_graph=use_graph)
    stop.record(eng.stream)
    stop.synchronize()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = start.elapsed_time(stop)
    return ms, wall


def copy_bandwidth(nbytes=2 << 30, reps=20):
    """Bandwidth of the library's plain copy kernel (read + write bytes)."""
    import ctypes
    import torch
    from pylabolt_amd import _native as nt
    lib = nt.lib()
    s = torch.cuda.Stream()
    sp = ctypes.c_void_p(s.cuda_stream)
    a = torch.empty(nbytes // 8, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    a.fill_(1.0)
    torch.cuda.synchronize()
    for _ in range(3):
        nt.check(lib.d2q9_copy(nt.ptr(b), nt.ptr(a), nbytes, sp), "copy")
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(s)
    for _ in range(reps):
        nt.check(lib.d2q9_copy(nt.ptr(b), nt.ptr(a), nbytes, sp), "copy")
    e1.record(s)
    e1.synchronize()
    ms = e0.elapsed_time(e1) / reps
    return 2.0 * nbytes / (ms * 1e-3) / 1e9


def cpu_baseline(sim, target_s=15.0, max_steps=200):
    """C++/OpenMP reference restatement on this host, bounded sample."""
    from oracle import ref_lib
    from tests.helpers import oracle_case
    case = oracle_case(sim)
    r = ref_lib.RefRunner(case)
    r.initialize()
    threads = ref_lib.max_threads()
    t0 = time.perf_counter()
    r.run(2)
    per = (time.perf_counter() - t0) / 2
    n = int(max(1, min(max_steps, target_s / max(per, 1e-9))))
    t0 = time.perf_counter()
    r.run(n)
    dt = time.perf_counter() - t0
    nx, ny = sim.mesh_dict["grid"]
    return {"value": nx * ny * n / dt / 1e6, "unit": "MLUPS",
            "cores": threads, "kind": "port",
            "sample": f"{n} steps of the same {nx}x{ny} workload after 2 "
                      f"warm-up steps, OpenMP threads={threads}",
            "host_cpu_count": os.cpu_count(), "host_cpu_model": cpu_model(),
            "omp_num_threads_env": os.environ.get("OMP_NUM_THREADS")}


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def adjoint_cost(steps=2000, reps=3):
    """Adjoint workload (Guo-forced channel 1024 x 256, FP64, 2000 steps):
    time of forward + backward over the plain forward of the same steps."""
    import torch
    from pylabolt_amd import cases
    from pylabolt_amd.adjoint import FlowRollout
    from pylabolt_amd.solvers.fluidLB import Solver
    sim, F, u0, target = cases.adjoint_channel(steps=steps)
    nx, ny = F.shape[:2]
    out = {"workload": f"adjoint channel BGK+Guo FP64 {nx}x{ny}, {steps} steps"}
    ro = {m: FlowRollout(sim, steps, checkpoint=m) for m in ("sqrt", "moments")}
    eng = ro["sqrt"].eng
    Ft = torch.tensor(F, device="cuda")
    ut = torch.tensor(u0, device="cuda")
    tt = torch.tensor(target, device="cuda")

    def timed(fn):
        best = None
        for _ in range(reps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            fn()
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
            best = dt if best is None else min(best, dt)
        return best

    # plain forward: one graph of the same init + N sweeps, moments stored
    # at the last step only, no checkpoints; its buffers are set up by a
    # rollout forward before the timer
    r0 = ro["sqrt"]
    r0.forward(Ft, ut)

    def issue_plain():
        from pylabolt_amd import _native as nt
        nt.check(eng.lib.d2q9_init_eq(eng._Lp, nt.ptr(r0.work[0]),
                                      nt.ptr(r0.work[1]), nt.ptr(r0.mom0),
                                      eng._stream_ptr), "d2q9_init_eq")
        prev = r0.work[0]
        for n in range(1, steps + 1):
            dst = r0.work[n % 2]
            r0._sweep(prev, dst, r0.mscratch[(n - 1) % 2], r0.mscratch[n % 2],
                      n == 1, n == steps)
            prev = dst
    g_plain = r0._capture(issue_plain)

    def plain():
        from pylabolt_amd import _native as nt
        nt.check(eng.lib.d2q9_graph_launch(g_plain, eng._stream_ptr),
                 "graph launch")
    t_plain = timed(plain)
    eng.lib.d2q9_graph_destroy(g_plain)
    out["plain_forward_ms"] = t_plain * 1e3
    out["plain_forward"] = ("one HIP graph: d2q9_init_eq + collide + "
                            f"{steps - 1} fused steps, no checkpoints")
    for m, r in ro.items():
        u = r.forward(Ft, ut)
        g = torch.zeros_like(u)
        g[-1, :, 0] = u[-1, :, 0] - tt
        t_f = timed(lambda: r.forward(Ft, ut))
        t_b = timed(lambda: r.backward(g))
        out[m] = {"forward_ms": t_f * 1e3, "backward_ms": t_b * 1e3,
                  "cost_ratio": (t_f + t_b) / t_plain,
                  "checkpoints": len(r.ck), "segment": r.K}
        r.close()
    return out


def batched_form(steps=5000, warmup=50, per_lattice_steps=100, n=256,
                 seed=0):
    """Batched workload: ``n`` seeded lattices (sizes uniform over the
    multiples of 32 in 64..1024 x 64..512, Re in [50, 400], even ids cavity
    BGK, odd ids Guo-forced channel MRT), FP32, one persistent launch per
    step replayed from HIP graphs, launch configuration from the
    autotuner.  Aggregate MLUPS = sum of cells * steps / t.  The same
    lattices advanced one launch per lattice per step are timed beside it
    over ``per_lattice_steps`` steps."""
    import torch
    from pylabolt_amd import cases
    from pylabolt_amd.batch import BatchAutotuner, BatchedLattices
    sims = cases.batch_cases(n=n, seed=seed, steps=steps)
    t0 = time.perf_counter()
    tuner = BatchAutotuner(reps=5)
    batch = BatchedLattices(sims, storage="f32", tuner=tuner, graph_steps=50)
    t_setup = time.perf_counter() - t0
    batch.advance(warmup)
    batch.synchronize()

    def timed(fn):
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        ev0.record(batch.stream)
        fn()
        ev1.record(batch.stream)
        ev1.synchronize()
        return ev0.elapsed_time(ev1) * 1e-3

    t = timed(lambda: batch.advance(steps, write_moments=True))
    cells = batch.cells
    mlups = cells * steps / t / 1e6
    achieved = mlups * 1e6 * BYTES_PER_CELL["f32"] / 1e9
    finite = all(bool(np.isfinite(e.moments_host()[1]).all())
                 for e in batch.engines[:8])
    mlups_pl = None
    if per_lattice_steps > 0:
        t_pl = timed(lambda: batch.advance_per_lattice(per_lattice_steps))
        mlups_pl = cells * per_lattice_steps / t_pl / 1e6
    out = {
        "workload": f"{n} lattices (seed {seed}), FP32, {cells} cells, "
                    f"{steps} steps",
        "lattices": n, "cells": cells, "steps": steps,
        "mlups": mlups, "us_per_step": t / steps * 1e6,
        "achieved_gbs": achieved, "frac": achieved / PEAK_GBS,
        "config": batch.config.as_dict(), "items_per_step": batch.n_items,
        "tuner_candidates": len(tuner.trials),
        "setup_and_tune_s": t_setup,
        "per_lattice_launches_mlups": mlups_pl,
        "speedup_vs_per_lattice_launches": (mlups / mlups_pl
                                            if mlups_pl else None),
        "velocity_finite": finite,
    }
    batch.close()
    return out


def load_pmc(workload, storage):
    path = os.path.join(ROOT, "profiles", f"pmc_{workload}_{storage}.json")
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    return None


def sweep_rate(engine, bpc, ms_step, cells):
    """HBM rate of the measured sweep: (achieved GB/s, single-step
    equivalent GB/s, two-step details or None).  A two-step launch moves
    bpc per deep cell and 2 bpc per rim cell for two updates."""
    single = bpc * cells / (ms_step * 1e-3) / 1e9
    if not getattr(engine, "two_step", False):
        return single, single, None
    deep = int(round(engine.deep_fraction * cells))
    rim = cells - deep
    launch_bytes = bpc * deep + 2 * bpc * rim
    achieved = launch_bytes / (2 * ms_step * 1e-3) / 1e9
    return achieved, single, {
        "launch": "d2q9_fused_step2 (k_deep2 + rim sweeps), two steps per "
                  "launch",
        "deep_cells": deep, "rim_cells": rim, "deep_fraction": deep / cells,
        "algorithmic_bytes_per_launch": launch_bytes,
        "launch_us": 2 * ms_step * 1e3}


def run_form(sim, storage, steps, warmup, tile, use_graph=True):
    """Time ``steps`` steps of ``sim`` in one storage form after ``warmup``
    steps.  Every form asks for the two-step sweep (FP64 and FP32 would
    take it by size anyway; 16-bit storage only on request)."""
    from pylabolt_amd.solvers.fluidLB import Solver
    solver = Solver(simulation=sim, storage=storage, tile=tile,
                    graph_steps=graph_steps_for(steps), verbose=False,
                    two_step=True)
    solver.set_backend()
    ms, wall = time_engine(solver, steps, warmup, use_graph)
    return solver, ms, wall


def timed_advance(engine, n):
    """Advance ``n`` steps (moments stored at the last); GPU seconds by
    HIP events on the engine's stream."""
    import torch
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record(engine.stream)
    engine.advance(n, write_moments=True)
    e1.record(engine.stream)
    e1.synchronize()
    return e0.elapsed_time(e1) * 1e-3


def startup_series(y, H, nu, g, t, terms=400):
    """Start-up of plane Poiseuille flow between walls at y = 0 and H
    (uniform body force g from rest)."""
    u_max = g * H * H / (8 * nu)
    u = 4 * u_max * (y / H) * (1 - y / H)
    for k in range(terms):
        m = 2 * k + 1
        u -= (32 * u_max / (np.pi ** 3 * m ** 3)) * np.sin(m * np.pi * y / H) \
            * np.exp(-(m * np.pi / H) ** 2 * nu * t)
    return u


def channel_leg(storages, steps=None, warmup=100):
    """Gravity-driven channel 4096 x 1024, MRT + second-order Guo, Re = 100
    at u_max = 0.05, from rest, all its steps (20,000) in every storage
    form, two-step sweeps from HIP graphs.  Per form: MLUPS and roofline of
    the whole run, velocity error against the FP64 run (relative L2 over
    all cells) and the x-averaged profile against the start-up series."""
    import torch
    from pylabolt_amd import cases
    from pylabolt_amd.solvers.fluidLB import Solver
    sim = cases.poiseuille_channel()
    nx, ny = sim.mesh_dict["grid"]
    steps = steps or sim.control_dict["end_time"]
    nu = sim.transport_dict["kin_visc"]
    g = sim.forcing_dict["gravity"][0]
    y = np.arange(1, ny + 1) - 0.5
    series = startup_series(y, ny, nu, g, steps)
    out = {"workload": f"poiseuille channel D2Q9 {model_name(sim)} "
                       f"{nx}x{ny}, {steps} steps from rest",
           "steps": steps}
    ref = None
    for storage in storages:
        s = Solver(simulation=sim, storage=storage, verbose=False,
                   graph_steps=graph_steps_for(steps), two_step=True)
        s.set_backend()
        eng = s.engine
        eng.capture_graph()
        t = timed_advance(eng, steps)
        ms_step = t * 1e3 / steps
        bpc = BYTES_PER_CELL[storage]
        achieved, single, two = sweep_rate(eng, bpc, ms_step, nx * ny)
        _, vel = eng.moments_host()
        u = vel.reshape(nx + 2, ny + 2, 2)[1:-1, 1:-1]
        prof = u[:, :, 0].mean(axis=0)
        rec = {"mlups": nx * ny / (ms_step * 1e-3) / 1e6,
               "us_per_step": ms_step * 1e3, "gbs": achieved,
               "frac": achieved / PEAK_GBS,
               "single_step_equivalent_gbs": single,
               "two_step": two is not None,
               "u_rel_l2_vs_startup_series": float(
                   np.linalg.norm(prof - series) / np.linalg.norm(series))}
        if ref is None:
            ref = u.copy()
        else:
            rec["u_rel_l2_vs_f64"] = float(np.linalg.norm(u - ref) /
                                           np.linalg.norm(ref))
            rec["u_rms_err_over_u_max"] = float(
                np.sqrt(((u - ref) ** 2).sum(-1).mean()) / 0.05)
            rec["speedup_vs_f64"] = out["f64"]["us_per_step"] / (ms_step * 1e3)
        out[storage] = rec
        log(f"channel {storage}: {ms_step * 1e3:.1f} us/step "
            f"{rec['mlups']:.0f} MLUPS ({rec['frac']:.3f} of peak), u vs "
            f"series {rec['u_rel_l2_vs_startup_series']:.2e}"
            + (f", vs f64 {rec['u_rel_l2_vs_f64']:.2e}"
               if "u_rel_l2_vs_f64" in rec else ""))
        del s, eng
        torch.cuda.empty_cache()
    return out


class ChildLeg:
    """One leg of the benchmark (``module:function`` returning a JSON-able
    dict) in a child process.  The child is started before this process
    touches the GPU and waits for a line on stdin; ``run`` sends it, waits
    for the child's JSON record and returns it, or an ``error`` record when
    the child fails (an abort there ends the child, not the benchmark
    line).  Closing the pipe without ``run`` lets the child exit unused."""

    def __init__(self, target):
        self.target = target
        self.proc = subprocess.Popen(
            [sys.executable, os.path.abspath(__file__), "--child-leg", target],
            stdin=subprocess.PIPE, stdout=subprocess.PIPE, text=True)

    def run(self, timeout=900):
        try:
            out, _ = self.proc.communicate("go\n", timeout=timeout)
        except subprocess.TimeoutExpired:
            self.proc.kill()
            self.proc.communicate()
            return {"error": f"{self.target}: no result in {timeout} s"}
        recs = [ln for ln in out.splitlines() if ln.startswith("{")]
        if self.proc.returncode != 0 or not recs:
            return {"error": f"{self.target}: child exited with status "
                             f"{self.proc.returncode}"}
        return json.loads(recs[-1])

    def close(self):
        if self.proc.poll() is None:
            self.proc.communicate("", timeout=60)


def child_leg(target, real_stdout):
    """Child side of ``ChildLeg``: wait for the go line, run the leg, write
    its record as one JSON line."""
    if sys.stdin.readline().strip() != "go":
        return
    mod, fn = target.split(":")
    out = getattr(importlib.import_module(mod), fn)()
    print(json.dumps(out), file=real_stdout, flush=True)


def main():
    # everything but the final JSON line goes to stderr
    real_stdout = sys.stdout
    sys.stdout = sys.stderr
    try:
        if len(sys.argv) == 3 and sys.argv[1] == "--child-leg":
            child_leg(sys.argv[2], real_stdout)
            return
        out = run(real_stdout)
    finally:
        sys.stdout = real_stdout
    if out is not None:
        print(json.dumps(out), flush=True)


def run(real_stdout):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--workload", default="cylinder")
    ap.add_argument("--nx", type=int, default=0)
    ap.add_argument("--ny", type=int, default=0)
    ap.add_argument("--storage", default="f64")
    ap.add_argument("--tile", default="64x4", help="tile rows: must be 4")
    ap.add_argument("--forms", default="f32,f16dev,bf16dev",
                    help="extra storage forms to time ('' for none)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-copy", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-channel", action="store_true")
    ap.add_argument("--long-steps", type=int, default=50000,
                    help="cylinder forms run on to this step count for "
                    "their velocity error (0: not)")
    ap.add_argument("--profile-run", action="store_true",
                    help="a short run for rocprofv3: main kernel + copy "
                    "(--workload batched: the batched sweep only)")
    ap.add_argument("--eager", action="store_true",
                    help="profile run without HIP graphs (for counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must "
                          "be 1 (multi-GPU sharding is a design note only)"}),
              file=real_stdout)
        sys.exit(2)
    # the adjoint leg's child starts before this process touches the GPU
    adj_leg = (ChildLeg("bench:adjoint_cost")
               if not (args.no_adjoint or args.profile_run) else None)
    try:
        return run_legs(args, adj_leg)
    finally:
        if adj_leg is not None:
            adj_leg.close()


def run_legs(args, adj_leg):
    import torch
    if not torch.cuda.is_available():
        raise SystemExit("bench.py needs a GPU")
    tile = tuple(int(v) for v in args.tile.split("x"))
    if args.profile_run and args.workload == "batched":
        bat = batched_form(args.steps, args.warmup, per_lattice_steps=0)
        bw = copy_bandwidth(1 << 30, reps=5)
        log(f"profile run batched: {bat['us_per_step']:.1f} us/step "
            f"{bat['mlups']:.0f} MLUPS config {bat['config']}, copy {bw:.0f} GB/s")
        return None
    sim = make_workload(args.workload, args.nx, args.ny,
                        args.steps + args.warmup)
    nx, ny = sim.mesh_dict["grid"]
    cells = nx * ny

    if args.profile_run:
        solver, ms, wall = run_form(sim, args.storage, args.steps,
                                    args.warmup, tile,
                                    use_graph=not args.eager)
        bw = copy_bandwidth(1 << 30, reps=5)
        log(f"profile run: {ms / args.steps * 1e3:.1f} us/step, copy {bw:.0f} GB/s")
        return None

    log(f"workload {args.workload} {nx}x{ny} storage={args.storage} "
        f"tile={tile} steps={args.steps} warmup={args.warmup} "
        f"graph chunk {graph_steps_for(args.steps)}")
    long_steps = args.long_steps if args.workload == "cylinder" else 0
    forms = {}
    ref_vel = None
    main = None
    storages = [args.storage] + [f for f in args.forms.split(",")
                                 if f and f != args.storage]
    for form in storages:
        s2, ms2, wall2 = run_form(sim, form, args.steps, args.warmup, tile)
        st2 = ms2 / args.steps
        m2 = cells / (st2 * 1e-3) / 1e6
        a2, single2, two2 = sweep_rate(s2.engine, BYTES_PER_CELL[form], st2,
                                       cells)
        rec = {"mlups": m2, "us_per_step": st2 * 1e3, "gbs": a2,
               "frac": a2 / PEAK_GBS, "two_step": two2 is not None,
               "single_step_equivalent_gbs": single2}
        if main is None:
            main = (ms2, wall2, st2, m2, a2, single2, two2)
            log(f"{form}: {st2 * 1e3:.1f} us/step  {m2:.0f} MLUPS  "
                f"{a2:.0f} GB/s ({a2 / PEAK_GBS:.3f} of peak), wall "
                f"{wall2:.2f}s")
        else:
            rec["speedup_vs_" + args.storage] = m2 / main[3]
            log(f"{form}: {st2 * 1e3:.1f} us/step {m2:.0f} MLUPS {a2:.0f} "
                f"GB/s ({a2 / PEAK_GBS:.3f} of peak)")
        if long_steps:
            eng = s2.engine
            n_long = long_steps - eng.time_step
            t_long = timed_advance(eng, n_long)
            _, vel2 = eng.moments_host()
            fluid = eng.tables.fluid_pad.ravel()
            rec["long_run"] = {"steps": long_steps,
                               "us_per_step": t_long * 1e6 / n_long}
            if ref_vel is None:
                ref_vel = vel2
                rec["long_run"]["u_rms"] = float(np.sqrt(
                    (vel2[fluid] ** 2).sum(axis=-1).mean()))
            else:
                du = vel2[fluid] - ref_vel[fluid]
                rec["long_run"]["u_rel_l2_vs_" + args.storage] = float(
                    np.sqrt((du ** 2).sum() / max(
                        (ref_vel[fluid] ** 2).sum(), 1e-300)))
                rec["long_run"]["u_rms_err_over_u_ref"] = float(
                    np.sqrt((du ** 2).sum(axis=-1).mean())) / U_REF
                log(f"{form} after {long_steps} steps: u rel L2 vs "
                    f"{args.storage} "
                    f"{rec['long_run']['u_rel_l2_vs_' + args.storage]:.2e}")
        forms[form] = rec
        del s2
        torch.cuda.empty_cache()
    ms, wall, ms_step, mlups, achieved, single_equiv, two = main
    bpc = BYTES_PER_CELL[args.storage]

    chan = None
    if not args.no_channel:
        chan = channel_leg(storages)
        torch.cuda.empty_cache()

    bat = None
    if not args.no_batched:
        bat = batched_form()
        log(f"batched: {bat['mlups']:.0f} MLUPS ({bat['frac']:.1%} of peak) "
            f"config {bat['config']}, per-lattice launches "
            f"{bat['per_lattice_launches_mlups']:.0f} MLUPS")
        torch.cuda.empty_cache()
    adj = None
    if adj_leg is not None:
        torch.cuda.synchronize()
        torch.cuda.empty_cache()
        adj = adj_leg.run()
        if "error" in adj:
            log(f"adjoint: {adj['error']}")
        else:
            log(f"adjoint: plain fwd {adj['plain_forward_ms']:.1f} ms, sqrt "
                f"ratio {adj['sqrt']['cost_ratio']:.2f}, moments ratio "
                f"{adj['moments']['cost_ratio']:.2f}")
    copy_gbs = None if args.no_copy else copy_bandwidth()
    if copy_gbs:
        log(f"copy kernel: {copy_gbs:.0f} GB/s")
    pmc = load_pmc(args.workload, args.storage)
    traffic = pmc["hbm_bytes_per_launch"] if pmc else None
    if pmc and two is not None and pmc.get("launch") != "two-step":
        traffic = None     # counters of another launch shape
    traffic_source = (f"profiles/pmc_{args.workload}_{args.storage}.json: "
                      "rocprofv3 FETCH_SIZE and WRITE_SIZE of the same "
                      "launch, scaled by a copy kernel's counters "
                      "(not measured in this run)" if traffic else None)
    cpu = None
    if not args.no_cpu:
        log("cpu baseline ...")
        cpu = cpu_baseline(sim)
        log(f"cpu: {cpu['value']:.1f} MLUPS on {cpu['cores']} threads")

    out = {
        "metric": "MLUPS + % HBM roofline (FP64 vs FP32/FP16-store); "
                  "adjoint/forward cost ratio",
        "value": mlups,
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": DTYPES[args.storage],
        "data": "synthetic (seeded case file, fluid at rest, rho=1)",
        "config": {"workload": f"{args.workload} D2Q9 {model_name(sim)} "
                               f"{nx}x{ny}",
                   "storage": args.storage, "cells": cells,
                   "tile": f"{tile[0]}x{tile[1]}", "parallelism": "1 GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / PEAK_GBS,
                     "traffic": traffic,
                     "traffic_source": traffic_source,
                     "bytes_per_cell_update": (
                         two["algorithmic_bytes_per_launch"] / (2 * cells)
                         if two else bpc),
                     "single_step_bytes_per_cell_update": bpc,
                     "single_step_equivalent_gbs": single_equiv},
        "cpu_baseline": cpu,
        "forms": forms,
        "channel": chan,
        "batched": bat,
        "adjoint": adj,
        "copy_kernel_gbs": copy_gbs,
    }
    if two is not None:
        out["roofline"].update(two)
    if copy_gbs:
        out["roofline"]["frac_of_copy"] = achieved / copy_gbs
    return out


if __name__ == "__main__":
    main()
