"""Many independent lattices advanced together: one persistent launch per step.

The reference runs one case per ``Solver`` (pylabolt/solvers/fluidLB.py:
107-395); a sweep over sizes and Reynolds numbers is that loop run once per
case.  ``BatchedLattices`` keeps one engine per case -- each with its own
tables, collision model, forcing and boundaries -- and advances all of them
by one fused step in a single launch of ``d2q9_batched_step``: the step's
work is a list of items (bulk tiles, general tiles and chunks of fix-up
cells of every lattice) cut into eight queues, one per XCD (small lattices
whole, large ones in row-ordered pieces), and a fixed grid of workgroups
walks them statically or claims them with atomic counters.
Full chunks of steps replay a captured HIP graph.

``BatchAutotuner`` picks the tile shape (rows per workgroup), the register
budget of the kernel build (waves per SIMD), the queue policy, the grid and
the claim size by timing candidates on the batch itself, and caches the
winner per size class of the batch.
"""
import ctypes
import hashlib
import json
import math
import os
import time

import numpy as np
import torch

from pylabolt_amd import _native as nt
from pylabolt_amd.engine import ORIGIN

N_XCD = 8
POLICY_STATIC, POLICY_DYNAMIC = 0, 1


def work_items(engines, tile_i):
    """Per-lattice work items ``(lattice, code)``.

    code >= 0: position in the lattice's tile list; code < 0: fix-up chunk
    ``-code - 1`` of ``64 * tile_i`` cells.  Within a lattice the tiles go
    in row-band order and each fix-up chunk right after the band holding
    the last row it touches: its cells' cache lines were just fetched by
    those tiles and are still in the XCD's L2.
    """
    per = []
    for l, e in enumerate(engines):
        t = e.tables
        wg = 64 * tile_i
        nchunk = -(-t.n_fix // wg)
        ntiles = t.n_bulk + t.n_general
        band = np.asarray(t.tile_list[:ntiles]) // t.ntj
        by_band = np.argsort(band, kind="stable")
        # last interior row (1-based) of each fix-up chunk; the list is
        # sorted by row
        rows = (np.asarray(t.fix_cells[:t.n_fix], dtype=np.int64) - ORIGIN) \
            // e.pitch
        last = [int(rows[min((c + 1) * wg, t.n_fix) - 1])
                for c in range(nchunk)]
        codes, ci, k = [], 0, 0
        for b in range(t.nti):
            while k < ntiles and band[by_band[k]] == b:
                codes.append(int(by_band[k]))
                k += 1
            # rows up to (b + 1) * tile_i are swept; the chunk also reads
            # one row beyond its last
            while ci < nchunk and last[ci] + 1 <= (b + 1) * tile_i:
                codes.append(-(ci + 1))
                ci += 1
        codes += [int(p) for p in by_band[k:]]
        codes += [-(c + 1) for c in range(ci, nchunk)]
        per.append(np.stack([np.full(len(codes), l, dtype=np.int32),
                             np.asarray(codes, dtype=np.int32)], axis=1)
                   if codes else np.zeros((0, 2), dtype=np.int32))
    return per


def pack_queues(engines, per_items, n_queues=N_XCD, split=4):
    """Cut the lattices' item lists into pieces and assign the pieces to
    the XCD queues, largest first onto the least-loaded queue (LPT); each
    queue is the concatenation of its pieces.

    A lattice whose item list is longer than 1 / (n_queues * split) of the
    whole batch is cut into contiguous pieces of about that size (its rows
    stay in order inside a piece); smaller lattices stay whole, so their
    tiles share one XCD's L2.  The cost of a piece is its item count.

    Returns (items [n, 2] int32, bounds [n_queues + 1] int32, owner) with
    owner[q] the list of (lattice, first, stop) pieces of queue q.
    """
    total = sum(int(a.shape[0]) for a in per_items)
    cap = max(1, -(-total // (n_queues * split)))
    pieces = []
    for l, a in enumerate(per_items):
        n = int(a.shape[0])
        k = max(1, -(-n // cap))
        edges = [n * p // k for p in range(k + 1)]
        pieces += [(l, edges[p], edges[p + 1]) for p in range(k)
                   if edges[p + 1] > edges[p]]
    order = sorted(range(len(pieces)),
                   key=lambda p: (-(pieces[p][2] - pieces[p][1]), p))
    load = [0] * n_queues
    owner = [[] for _ in range(n_queues)]
    for p in order:
        q = min(range(n_queues), key=lambda x: (load[x], x))
        owner[q].append(pieces[p])
        load[q] += pieces[p][2] - pieces[p][1]
    chunks, bounds = [], [0]
    for q in range(n_queues):
        owner[q].sort()
        for l, a, b in owner[q]:
            chunks.append(per_items[l][a:b])
        bounds.append(bounds[-1] + sum(b - a for _, a, b in owner[q]))
    items = (np.concatenate(chunks) if chunks else
             np.zeros((0, 2), dtype=np.int32))
    return (np.ascontiguousarray(items, dtype=np.int32),
            np.asarray(bounds, dtype=np.int32), owner)


class BatchConfig:
    """Launch configuration: rows per tile, queue policy, grid, claim size,
    register budget (waves per SIMD; 0 = the compiler's allocation)."""

    def __init__(self, tile_i=4, policy=POLICY_DYNAMIC, grid=None, chunk=1,
                 waves=0):
        self.tile_i, self.policy = int(tile_i), int(policy)
        self.grid, self.chunk, self.waves = grid, int(chunk), int(waves)

    def as_dict(self):
        return {"tile_i": self.tile_i,
                "policy": "dynamic" if self.policy else "static",
                "grid": self.grid, "chunk": self.chunk, "waves": self.waves}

    @classmethod
    def from_dict(cls, d):
        return cls(d["tile_i"], POLICY_DYNAMIC if d["policy"] == "dynamic"
                   else POLICY_STATIC, d["grid"], d["chunk"],
                   d.get("waves", 0))

    def __repr__(self):
        return f"BatchConfig({self.as_dict()})"


class BatchedLattices:
    """Lockstep advance of many lattices through one launch per step.

    ``simulations``: case namespaces (the reference's ``simulation.py``
    dicts), each built into its own engine on one shared stream; they must
    share the storage form (FP64 or FP32).  ``config``: a BatchConfig, or
    None for the autotuner's choice (``tuner``, default a
    BatchAutotuner with its cache).
    """

    def __init__(self, simulations, storage="f32", config=None, tuner=None,
                 graph_steps=16, device=None):
        if storage not in ("f64", "f32"):
            raise ValueError("batched lattices use FP64 or FP32 storage")
        from pylabolt_amd.solvers.fluidLB import Solver
        self.lib = nt.lib()
        self.device = (torch.device(device) if device is not None else
                       torch.device("cuda", torch.cuda.current_device()))
        self.stream = torch.cuda.Stream(device=self.device)
        self._sp = ctypes.c_void_p(self.stream.cuda_stream)
        self.storage = storage
        self.solvers = []
        for sim in simulations:
            # the batched launches step every lattice one step at a time:
            # no two-step scratch lattice
            s = Solver(simulation=sim, storage=storage, verbose=False,
                       stream=self.stream, two_step=False)
            s.set_backend(verbose=False)
            self.solvers.append(s)
        self.engines = [s.engine for s in self.solvers]
        if not self.engines:
            raise ValueError("a batch needs at least one lattice")
        self.graph_steps = max(2, int(graph_steps) // 2 * 2)
        self.parity = 0
        self._graphs = {}
        with torch.cuda.stream(self.stream):
            self.counters = torch.zeros(2 * N_XCD, dtype=torch.int32,
                                        device=self.device)
            self.entries = torch.zeros(
                len(self.engines) * ctypes.sizeof(nt.BatchEntry),
                dtype=torch.uint8, device=self.device)
        self._entry_key = None
        self.cells = sum(e.nx * e.ny for e in self.engines)
        self.config = None
        if config is None:
            tuner = tuner if tuner is not None else BatchAutotuner()
            config = tuner.config_for(self)
        self.set_config(config)

    # ---- configuration ----------------------------------------------------
    def resident_grid(self, tile_i, waves=0):
        """Workgroups that fit on the GPU at once: 64 x tile_i threads
        each, ``waves`` per SIMD (four SIMDs per CU) for the register-capped
        builds, four for the default build."""
        cus = torch.cuda.get_device_properties(self.device).multi_processor_count
        per_cu = max(1, 4 * (waves or 4) // tile_i)
        return int(cus * per_cu)

    def set_config(self, config):
        cfg = BatchConfig(config.tile_i, config.policy, config.grid,
                          config.chunk, config.waves)
        if cfg.grid is None:
            cfg.grid = self.resident_grid(cfg.tile_i, cfg.waves)
        cfg.grid = max(N_XCD, int(cfg.grid) // N_XCD * N_XCD)
        if self.config is None or cfg.tile_i != self.config.tile_i:
            self.stream.synchronize()
            for e in self.engines:
                e.set_tile_rows(cfg.tile_i)
            items, bounds, self.owner = pack_queues(
                self.engines, work_items(self.engines, cfg.tile_i))
            with torch.cuda.stream(self.stream):
                self.items = torch.from_numpy(items.reshape(-1)).to(
                    self.device)
                self.bounds = torch.from_numpy(bounds).to(self.device)
            self.n_items = int(items.shape[0])
            self.has_general = int(any(e.tables.n_general > 0
                                       for e in self.engines))
            self.queue_sizes = np.diff(bounds)
            self._entry_key = None
        self._drop_graphs()
        self.config = cfg

    # ---- state ------------------------------------------------------------
    def _state_key(self):
        return tuple((e.i_pop ^ self.parity, e.i_mom ^ self.parity,
                      e.tile_list.data_ptr()) for e in self.engines)

    def _sync_entries(self):
        """(Re)write the entry table so that entry l's f[parity] is engine
        l's current population buffer.  Captured graphs read the table by
        pointer, so rewriting it in place keeps them valid."""
        key = self._state_key()
        if key == self._entry_key:
            return
        arr = (nt.BatchEntry * len(self.engines))()
        p = self.parity
        for l, e in enumerate(self.engines):
            arr[l].L = e.L
            arr[l].f[p] = e.f[e.i_pop].data_ptr()
            arr[l].f[1 - p] = e.f[1 - e.i_pop].data_ptr()
            arr[l].mom[p] = e.mom[e.i_mom].data_ptr()
            arr[l].mom[1 - p] = e.mom[1 - e.i_mom].data_ptr()
        host = torch.frombuffer(bytearray(bytes(arr)), dtype=torch.uint8)
        with torch.cuda.stream(self.stream):
            self.entries.copy_(host)
        self.stream.synchronize()
        self._entry_key = key

    def _launch(self, parity, write_moments):
        c = self.config
        return self.lib.d2q9_batched_step(
            nt.ptr(self.entries), nt.ptr(self.items), nt.ptr(self.bounds),
            nt.ptr(self.counters), int(parity),
            self.engines[0].storage, c.tile_i, c.grid, c.policy, c.chunk,
            c.waves, self.has_general, int(write_moments), self._sp)

    def _after_steps(self, k, write_moments):
        flip = k & 1
        for e in self.engines:
            e.i_pop ^= flip
            e.i_mom ^= flip
            e.i_pop_new = None
            e.time_step += k
            e.mom_full = bool(write_moments)
        self.parity ^= flip

    def _drop_graphs(self):
        for g in self._graphs.values():
            self.lib.d2q9_graph_destroy(g)
        self._graphs = {}

    def close(self):
        self._drop_graphs()
        for e in self.engines:
            e.close()

    def __del__(self):
        try:
            self._drop_graphs()
        except Exception:
            pass

    def _graph(self):
        g = self._graphs.get(self.parity)
        if g is not None:
            return g
        nt.check(self.lib.d2q9_graph_begin(self._sp), "graph begin")
        err = None
        p = self.parity
        for _ in range(self.graph_steps):
            if self._launch(p, 0) != 0 and err is None:
                err = self.lib.d2q9_last_error().decode()
            p ^= 1
        handle = ctypes.c_void_p()
        rc = self.lib.d2q9_graph_end(self._sp, ctypes.byref(handle))
        if err is not None:
            raise nt.NativeError("batched graph capture: " + err)
        nt.check(rc, "graph end")
        self._graphs[self.parity] = handle
        return handle

    # ---- stepping ---------------------------------------------------------
    @property
    def time_step(self):
        return self.engines[0].time_step

    def advance(self, n, write_moments=True, use_graph=True):
        """Advance every lattice ``n`` steps; the last one stores rho/u on
        every fluid cell when ``write_moments``."""
        n = int(n)
        if n <= 0:
            return
        t0 = self.engines[0].time_step
        if any(e.time_step != t0 for e in self.engines):
            raise RuntimeError("batched lattices must advance in lockstep")
        if any(e.i_pop_new is not None for e in self.engines):
            # first step after initialisation: collide-only, per lattice
            for e in self.engines:
                e.step(write_moments and n == 1)
            n -= 1
        if n == 0:
            return
        self._sync_entries()
        if use_graph:
            while n - 1 >= self.graph_steps:
                g = self._graph()
                nt.check(self.lib.d2q9_graph_launch(g, self._sp),
                         "graph launch")
                self._after_steps(self.graph_steps, False)
                n -= self.graph_steps
        while n > 0:
            last = write_moments and n == 1
            nt.check(self._launch(self.parity, last), "d2q9_batched_step")
            self._after_steps(1, last)
            n -= 1

    def advance_per_lattice(self, n, write_moments=True):
        """Advance every lattice ``n`` steps with its own launches (one
        sweep per lattice per step, graphs per lattice): the
        one-case-at-a-time schedule, kept as the comparison point."""
        for e in self.engines:
            e.advance(n, write_moments=write_moments)
        self._entry_key = None

    def synchronize(self):
        self.stream.synchronize()

    def timed_repeat(self, reps):
        """Seconds per step of ``reps`` launches of the same step (parity
        unchanged, so the state stays as it is: each launch rewrites the
        spare buffers from the current ones).  The claim counters are
        cleared between launches.  A lattice whose first (collide-only)
        step is still due keeps its streamed populations in the spare
        buffer; they are saved and put back."""
        self._sync_entries()
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        with torch.cuda.stream(self.stream):
            saved = [(e, e.f[e.i_pop_new].clone()) for e in self.engines
                     if e.i_pop_new is not None]
            self.counters.zero_()
            nt.check(self._launch(self.parity, 0), "d2q9_batched_step")
            self.counters.zero_()
            ev0.record(self.stream)
            for _ in range(reps):
                nt.check(self._launch(self.parity, 0), "d2q9_batched_step")
                self.counters.zero_()
            ev1.record(self.stream)
            for e, buf in saved:
                e.f[e.i_pop_new].copy_(buf)
        ev1.synchronize()
        self.stream.synchronize()
        return ev0.elapsed_time(ev1) * 1e-3 / reps


class BatchAutotuner:
    """Chooses tile shape, register budget, queue policy, grid and claim
    size per size class.

    A batch's size class is its storage form plus the count of lattices per
    (log2 of nx, number of row blocks along y, collision model, forcing)
    class; the winner is cached in memory and, when ``cache_path`` is set,
    in a JSON file.
    """

    TILE_ROWS = (4, 2, 1)
    WAVES = (0, 6, 8)
    GRID_FACTORS = (1, 2)
    CHUNKS = (1, 4)

    def __init__(self, cache_path=None, reps=5, verbose=False):
        self.cache_path = cache_path
        self.reps = int(reps)
        self.verbose = verbose
        self.cache = {}
        self.trials = []
        if cache_path and os.path.exists(cache_path):
            with open(cache_path) as fh:
                self.cache = json.load(fh)

    @staticmethod
    def signature(batch):
        hist = {}
        for e in batch.engines:
            key = (int(math.ceil(math.log2(max(e.nx, 1)))),
                   -(-e.ny // e.B), e.L.collision, e.L.forcing)
            hist[key] = hist.get(key, 0) + 1
        text = json.dumps([batch.storage, sorted(
            [list(k) + [v] for k, v in hist.items()])])
        return hashlib.sha1(text.encode()).hexdigest()[:16]

    def candidates(self, batch):
        out = []
        for ti in self.TILE_ROWS:
            for wv in self.WAVES:
                base = batch.resident_grid(ti, wv)
                for gf in self.GRID_FACTORS:
                    out.append(BatchConfig(ti, POLICY_STATIC, base * gf, 1,
                                           wv))
                    for ch in self.CHUNKS:
                        out.append(BatchConfig(ti, POLICY_DYNAMIC, base * gf,
                                               ch, wv))
        return out

    def tune(self, batch):
        """Time every candidate on the batch (state unchanged) and return
        the fastest."""
        best, best_t = None, float("inf")
        self.trials = []
        for cfg in self.candidates(batch):
            batch.set_config(cfg)
            t = batch.timed_repeat(self.reps)
            self.trials.append((cfg.as_dict(), t))
            if self.verbose:
                print(f"[tune] {cfg.as_dict()} {t * 1e3:.3f} ms", flush=True)
            if t < best_t:
                best, best_t = cfg, t
        return best, best_t

    def config_for(self, batch):
        sig = self.signature(batch)
        hit = self.cache.get(sig)
        if hit is not None:
            return BatchConfig.from_dict(hit["config"])
        # the first candidate gives the batch tables to time with
        batch.set_config(self.candidates(batch)[0])
        best, t = self.tune(batch)
        self.cache[sig] = {"config": best.as_dict(), "seconds_per_step": t}
        if self.cache_path:
            os.makedirs(os.path.dirname(os.path.abspath(self.cache_path)),
                        exist_ok=True)
            with open(self.cache_path, "w") as fh:
                json.dump(self.cache, fh, indent=1)
        return best
