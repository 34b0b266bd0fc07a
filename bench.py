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

Also reported: ``forms``, FP32 and 16-bit-storage runs of the same
workload with their roofline fractions and velocity error against the
FP64 run (relative L2 over the fluid cells, and RMS error over the case's
reference speed u_ref = 0.05); ``batched``, 256 seeded lattices of uneven size advanced by one
persistent launch per step; ``adjoint``, the forward + backward cost of the
adjoint channel over its plain forward; and a plain copy kernel's bandwidth
on the same device.
"""
import argparse
import json
import os
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


def time_engine(solver, steps, warmup, use_graph=True):
    import torch
    eng = solver.engine
    eng.advance(warmup, write_moments=False, use_graph=use_graph)
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
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
                      f"warm-up steps, OpenMP threads={threads}"}


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
    # plain forward: the same lattice, force field, graph-replayed steps
    from pylabolt_amd.engine import D2Q9Engine, elements_from_boundary
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

    def plain():
        eng.initialize()
        eng.advance(steps, write_moments=True)
        eng.synchronize()
    t_plain = timed(plain)
    out["plain_forward_ms"] = t_plain * 1e3
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
    from pylabolt_amd.solvers.fluidLB import Solver
    solver = Solver(simulation=sim, storage=storage, tile=tile,
                    graph_steps=32, verbose=False)
    solver.set_backend()
    ms, wall = time_engine(solver, steps, warmup, use_graph)
    return solver, ms, wall


def main():
    # everything but the final JSON line goes to stderr
    real_stdout = sys.stdout
    sys.stdout = sys.stderr
    try:
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
    import torch
    if not torch.cuda.is_available():
        raise SystemExit("bench.py needs a GPU")
    tile = tuple(int(v) for v in args.tile.split("x"))
    if args.profile_run and args.workload == "batched":
        bat = batched_form(args.steps, args.warmup, per_lattice_steps=0)
        log(f"profile run batched: {bat['us_per_step']:.1f} us/step "
            f"{bat['mlups']:.0f} MLUPS config {bat['config']}")
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
        f"tile={tile} steps={args.steps} warmup={args.warmup}")
    solver, ms, wall = run_form(sim, args.storage, args.steps, args.warmup,
                                tile)
    ms_step = ms / args.steps
    mlups = cells / (ms_step * 1e-3) / 1e6
    bpc = BYTES_PER_CELL[args.storage]
    achieved, single_equiv, two = sweep_rate(solver.engine, bpc, ms_step,
                                             cells)
    log(f"{args.storage}: {ms_step * 1e3:.1f} us/step  {mlups:.0f} MLUPS  "
        f"{achieved:.0f} GB/s ({achieved / PEAK_GBS:.3f} of peak), wall "
        f"{wall:.2f}s")
    ref_rho, ref_vel = None, None
    solver.engine.advance(1, write_moments=True)
    ref_rho, ref_vel = solver.engine.moments_host()
    fluid = solver.engine.tables.fluid_pad.ravel()
    steps_done = solver.engine.time_step
    del solver
    torch.cuda.empty_cache()

    forms = {args.storage: {"mlups": mlups, "us_per_step": ms_step * 1e3,
                            "gbs": achieved, "frac": achieved / PEAK_GBS}}
    for form in [f for f in args.forms.split(",") if f and f != args.storage]:
        s2, ms2, _ = run_form(sim, form, args.steps, args.warmup, tile)
        st2 = ms2 / args.steps
        m2 = cells / (st2 * 1e-3) / 1e6
        a2, single2, two2 = sweep_rate(s2.engine, BYTES_PER_CELL[form], st2,
                                       cells)
        s2.engine.advance(steps_done - s2.engine.time_step, write_moments=True)
        _, vel2 = s2.engine.moments_host()
        du = vel2[fluid] - ref_vel[fluid]
        l2 = float(np.sqrt((du ** 2).sum() / max((ref_vel[fluid] ** 2).sum(),
                                                 1e-300)))
        rms_ref = float(np.sqrt((du ** 2).sum(axis=-1).mean())) / U_REF
        forms[form] = {"mlups": m2, "us_per_step": st2 * 1e3, "gbs": a2,
                       "frac": a2 / PEAK_GBS,
                       "two_step": two2 is not None,
                       "single_step_equivalent_gbs": single2,
                       "speedup_vs_" + args.storage: m2 / mlups,
                       "u_rel_l2_vs_" + args.storage: l2,
                       "u_rms_err_over_u_ref": rms_ref,
                       "u_rms_" + args.storage: float(np.sqrt(
                           (ref_vel[fluid] ** 2).sum(axis=-1).mean())),
                       "steps_compared": steps_done}
        log(f"{form}: {st2 * 1e3:.1f} us/step {m2:.0f} MLUPS {a2:.0f} GB/s "
            f"u L2 rel err {l2:.2e}, rms err / u_ref {rms_ref:.2e}")
        del s2
        torch.cuda.empty_cache()

    bat = None
    if not args.no_batched:
        bat = batched_form()
        log(f"batched: {bat['mlups']:.0f} MLUPS ({bat['frac']:.1%} of peak) "
            f"config {bat['config']}, per-lattice launches "
            f"{bat['per_lattice_launches_mlups']:.0f} MLUPS")
        torch.cuda.empty_cache()
    adj = None
    if not args.no_adjoint:
        adj = adjoint_cost()
        log(f"adjoint: plain fwd {adj['plain_forward_ms']:.1f} ms, sqrt "
            f"ratio {adj['sqrt']['cost_ratio']:.2f}, moments ratio "
            f"{adj['moments']['cost_ratio']:.2f}")
        torch.cuda.empty_cache()
    copy_gbs = None if args.no_copy else copy_bandwidth()
    if copy_gbs:
        log(f"copy kernel: {copy_gbs:.0f} GB/s")
    pmc = load_pmc(args.workload, args.storage)
    traffic = pmc["hbm_bytes_per_launch"] if pmc else None
    if pmc and two is not None and pmc.get("launch") != "two-step":
        traffic = None     # counters of another launch shape
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
                     "bytes_per_cell_update": (
                         two["algorithmic_bytes_per_launch"] / (2 * cells)
                         if two else bpc),
                     "single_step_bytes_per_cell_update": bpc,
                     "single_step_equivalent_gbs": single_equiv},
        "cpu_baseline": cpu,
        "forms": forms,
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
