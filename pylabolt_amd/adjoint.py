"""Gradients of a flow loss through N FP64 time steps, by a hand-written
adjoint (reverse-time collide-stream) kernel.

``FlowRollout`` runs the reference time step (BGK, Guo forcing) from the
initial velocity ``u0`` under a fixed per-cell body-force field ``F`` and
returns the velocity field of the last step; its ``backward`` takes
dL/du_N and returns dL/dF and dL/du0.  ``LBMFlow`` wraps the pair as a
``torch.autograd.Function``.

Memory/compute schedule (``checkpoint``):

* ``"sqrt"`` (default): the forward pass stores the post-collision
  populations every K = ceil(sqrt(N)) steps -- the step that produces a
  checkpoint writes straight into its slot, so checkpoints cost no extra
  traffic; the backward pass walks the segments last to first, re-running
  each segment's K forward steps to store their rho/u (24 B per cell per
  step), then running the K adjoint steps.  Memory: ~2 sqrt(N) lattices.
* ``"moments"``: the forward pass stores rho/u of every step (24 B per cell
  per step, 12.6 GB for 2000 steps of 1024 x 256) and the backward pass
  recomputes nothing.

The adjoint of a step needs the forward state only through that step's
rho/u (the collision is a function of the moments and the populations
enter it linearly), which is what makes the moment store sufficient.

Supported cases: BGK with Guo forcing (linear or second order), a
per-cell force field, bounce-back / periodic boundaries and static solids at
rest -- for these, streaming permutes the populations and its transpose is
an exact scatter.  The constructor rejects anything else.
"""
import ctypes
import math

import numpy as np
import torch

from pylabolt_amd import _native as nt
from pylabolt_amd.engine import (ORIGIN, D2Q9Engine, elements_from_boundary,
                                 padded_view)


class FlowRollout:
    """N-step FP64 rollout with an adjoint.

    ``simulation`` is a case namespace (reference dict API).  Its collision
    model must be BGK; the force field given to ``forward`` replaces
    ``rho * gravity``.
    """

    def __init__(self, simulation, steps, checkpoint="sqrt", device=None):
        from pylabolt_amd.base.collision_operator import CollisionOperator
        from pylabolt_amd.base.state import State
        from pylabolt_amd.solvers.fluidLB import FluidLB
        self.steps = int(steps)
        if self.steps < 1:
            raise ValueError("steps must be >= 1")
        if checkpoint not in ("sqrt", "moments"):
            raise ValueError("checkpoint must be 'sqrt' or 'moments'")
        self.mode = checkpoint
        st = State(simulation)
        co = CollisionOperator(simulation, FluidLB(), st, verbose=False)
        if co.collision_fluid != "BGK":
            raise ValueError("the adjoint supports the BGK collision model")
        if co.forcing_fluid is None:
            raise ValueError("the adjoint needs a Guo forcing model: the "
                             "force field acts through its source term")
        for e in st.boundary.boundary_elements:
            if e.type_fluid not in ("bounce_back", "periodic"):
                raise ValueError("the adjoint supports bounce_back and periodic "
                                 "boundaries, not " + e.type_fluid)
        f = st.fields
        if np.any(f.velocity[f.solid] != 0):
            raise ValueError("the adjoint needs solids at rest")
        self.state = st
        self.nx, self.ny = int(st.domain.Nx_rank), int(st.domain.Ny_rank)
        size = int(st.domain.size)
        lat = st.lattice
        self.eng = D2Q9Engine(
            self.nx, self.ny, f.solid, f.density, f.velocity,
            elements_from_boundary(st.boundary), st.boundary.x_periodic,
            st.boundary.y_periodic, "BGK", co.forcing_fluid, co.omega_fluid,
            (0.0, 0.0), lat.inv_cs_2, lat.inv_cs_4, lat.weights,
            st.control.float_min, storage="f64",
            force_field=np.zeros((size, 2)), device=device)
        self._check_permutation()
        eng = self.eng
        dev = eng.device
        P, L = eng.plane, eng.pop_len
        N = self.steps
        self.K = int(math.ceil(math.sqrt(N))) if self.mode == "sqrt" else N
        self.nseg = int(math.ceil(N / self.K))
        z = dict(dtype=torch.float64, device=dev)
        with torch.cuda.stream(eng.stream):
            nck = self.nseg if self.mode == "sqrt" else 1
            self.ck = [torch.zeros(L, **z) for _ in range(nck)]
            self.work = [torch.zeros(L, **z) for _ in range(2)]
            self.ring = [torch.zeros(3 * P, **z) for _ in range(self.K)]
            self.mscratch = [torch.zeros(3 * P, **z) for _ in range(2)]
            self.mom0 = torch.zeros(3 * P, **z)
            self.lam = [torch.zeros(L, **z) for _ in range(2)]
            self.grad_f = torch.zeros(2 * P, **z)
            self.grad_init = torch.zeros(3 * P, **z)
            self.seed = torch.zeros(2 * P, **z)
        self.fluid_pad = torch.from_numpy(eng.tables.fluid_pad).to(dev)
        self._rho0 = torch.from_numpy(
            np.asarray(f.density).reshape(self.nx + 2, self.ny + 2)).to(dev)
        self._fwd_graph = None
        self._bwd_graph = None
        self._last_moments = None

    # ---- checks --------------------------------------------------------
    def _check_permutation(self):
        """Streaming with the case's rules must map every fluid population
        to exactly one pre-collision slot (so that the transpose is a
        scatter without collisions)."""
        t = self.eng.tables
        nxp, nyp = self.nx + 2, self.ny + 2
        fluid = t.fluid_pad
        solid = (t.flags_pad & nt.CELL_SOLID) != 0
        CX = [0, 1, 0, -1, 0, 1, -1, -1, 1]
        CY = [0, 0, 1, 0, -1, 1, 1, -1, -1]
        INV = [0, 3, 4, 1, 2, 7, 8, 5, 6]
        hits = np.zeros((nxp, nyp, 9), dtype=np.int32)
        ii, jj = np.nonzero(fluid)
        rules = np.zeros((nxp, nyp, 9), dtype=np.int32)
        bc = (t.flags_pad & nt.CELL_BC) != 0
        bi, bj = np.nonzero(bc)
        rules[bi, bj, :] = t.bc_nodes[t.bc_slot_pad[bi, bj], :9] & 15
        for d in range(9):
            si, sj = ii - CX[d], jj - CY[d]
            bb = (rules[ii, jj, d] == nt.RULE_BB) | (solid[si, sj] & (d > 0))
            # bounce-back: the cell's own opposite population
            np.add.at(hits, (ii[bb], jj[bb], INV[d]), 1)
            si, sj = si[~bb], sj[~bb]
            if self.eng.L.x_periodic:
                si = np.where(si == 0, self.nx, np.where(si == self.nx + 1, 1, si))
            if self.eng.L.y_periodic:
                sj = np.where(sj == 0, self.ny, np.where(sj == self.ny + 1, 1, sj))
            ok = fluid[si, sj]
            np.add.at(hits, (si[ok], sj[ok], d), 1)
        if hits[fluid].min() != 1 or hits[fluid].max() != 1:
            raise ValueError("streaming with these boundaries does not permute "
                             "the populations (every non-periodic side needs "
                             "bounce-back); the adjoint scatter would be wrong")

    # ---- layout helpers ------------------------------------------------
    def _to_pitched(self, field, dst, q0=0):
        """(nx, ny, C) interior field -> planes q0.. of a pitched buffer."""
        nx, ny = self.nx, self.ny
        for q in range(field.shape[-1]):
            v = padded_view(dst, q0 + q, nx, ny, self.eng.pitch, self.eng.plane)
            v[1:-1, 1:-1] = field[:, :, q] * self.fluid_pad[1:-1, 1:-1]

    def _from_pitched(self, src, q0, count):
        nx, ny = self.nx, self.ny
        return torch.stack([
            padded_view(src, q0 + q, nx, ny, self.eng.pitch,
                        self.eng.plane)[1:-1, 1:-1].clone()
            for q in range(count)], dim=-1)

    # ---- kernel wrappers (all on the engine stream) ---------------------
    def _lib(self):
        return self.eng.lib

    def _sweep(self, src, dst, msrc, mdst, first, write_mom):
        e = self.eng
        fn = e.lib.d2q9_collide if first else e.lib.d2q9_fused_step
        rc = fn(e._Lp, nt.ptr(src), nt.ptr(dst), nt.ptr(msrc), nt.ptr(mdst),
                int(write_mom), e._stream_ptr)
        if rc:
            raise nt.NativeError(e.lib.d2q9_last_error().decode())

    def _forward_ops(self):
        """The forward pass as a list of sweeps (captured in one graph)."""
        N, K = self.steps, self.K
        e = self.eng
        ops = []
        # g_0 = feq(rho0, u0) lives in ck[0]; work[1] also receives it
        prev = self.ck[0]
        for n in range(1, N + 1):
            ckpt = self.mode == "sqrt" and n % K == 0 and n // K < self.nseg
            dst = self.ck[n // K] if ckpt else self.work[n % 2]
            if dst is prev:
                dst = self.work[(n + 1) % 2]
            if self.mode == "moments":
                msrc = self.ring[n - 2] if n >= 2 else self.mscratch[0]
                mdst = self.ring[n - 1]
            else:
                msrc = self.mscratch[(n - 1) % 2]
                mdst = self.mscratch[n % 2]
            ops.append((prev, dst, msrc, mdst, n == 1, n == N or
                        self.mode == "moments"))
            prev = dst
        self._last_moments = ops[-1][3]
        return ops

    def _capture(self, issue):
        e = self.eng
        nt.check(e.lib.d2q9_graph_begin(e._stream_ptr), "graph begin")
        err = None
        try:
            issue()
        except Exception as ex:  # end the capture before re-raising
            err = ex
        h = ctypes.c_void_p()
        rc = e.lib.d2q9_graph_end(e._stream_ptr, ctypes.byref(h))
        if err is not None:
            raise err
        nt.check(rc, "graph end")
        return h

    def _issue_forward(self):
        e = self.eng
        nt.check(e.lib.d2q9_init_eq(e._Lp, nt.ptr(self.ck[0]),
                                    nt.ptr(self.work[1]), nt.ptr(self.mom0),
                                    e._stream_ptr), "d2q9_init_eq")
        for op in self._forward_ops():
            self._sweep(*op)

    def _issue_backward(self):
        e = self.eng
        N, K = self.steps, self.K
        lam_cur, lam_nxt = self.lam[0], self.lam[1]
        for s in range(self.nseg - 1, -1, -1):
            n0, n1 = s * K + 1, min((s + 1) * K, N)
            if self.mode == "sqrt":
                # recompute the segment, storing rho/u of each step
                if s == 0:
                    nt.check(e.lib.d2q9_init_eq(
                        e._Lp, nt.ptr(self.work[0]), nt.ptr(self.work[1]),
                        nt.ptr(self.mom0), e._stream_ptr), "d2q9_init_eq")
                    prev = self.work[0]
                else:
                    prev = self.ck[s]
                for n in range(n0, n1 + 1):
                    dst = self.work[1] if prev is self.work[0] else self.work[0]
                    slot = n - n0
                    msrc = self.ring[slot - 1] if slot else self.mscratch[0]
                    self._sweep(prev, dst, msrc, self.ring[slot], n == 1, True)
                    prev = dst
            for n in range(n1, n0 - 1, -1):
                slot = (n - n0) if self.mode == "sqrt" else n - 1
                seed = self.seed if n == N else None
                rc = e.lib.d2q9_adjoint_step(
                    e._Lp, int(n >= 2), nt.ptr(lam_cur), nt.ptr(lam_nxt),
                    nt.ptr(self.ring[slot]), nt.ptr(seed),
                    nt.ptr(self.grad_f), e._stream_ptr)
                if rc:
                    raise nt.NativeError(e.lib.d2q9_last_error().decode())
                lam_cur, lam_nxt = lam_nxt, lam_cur
        nt.check(e.lib.d2q9_adjoint_init_grad(
            e._Lp, nt.ptr(lam_cur), nt.ptr(self.mom0), nt.ptr(self.grad_init),
            e._stream_ptr), "d2q9_adjoint_init_grad")

    # ---- public -----------------------------------------------------------
    def forward(self, force, u0):
        """force, u0: float64 (nx, ny, 2) tensors on the device.  Returns the
        velocity of the last step, (nx, ny, 2)."""
        e = self.eng
        torch.cuda.current_stream(e.device).synchronize()
        with torch.cuda.stream(e.stream):
            e.force_field.zero_()
            self._to_pitched(force.to(e.device, torch.float64), e.force_field)
            self.mom0.zero_()
            v = padded_view(self.mom0, 0, self.nx, self.ny, e.pitch, e.plane)
            v[:] = self._rho0
            self._to_pitched(u0.to(e.device, torch.float64), self.mom0, 1)
        if self._fwd_graph is None:
            self._fwd_graph = self._capture(self._issue_forward)
        nt.check(e.lib.d2q9_graph_launch(self._fwd_graph, e._stream_ptr),
                 "graph launch")
        with torch.cuda.stream(e.stream):
            u = self._from_pitched(self._last_moments, 1, 2)
        e.stream.synchronize()
        return u

    def backward(self, grad_u):
        """grad_u = dL/du_N, float64 (nx, ny, 2).  Returns (dL/dF, dL/du0),
        both (nx, ny, 2)."""
        e = self.eng
        torch.cuda.current_stream(e.device).synchronize()
        with torch.cuda.stream(e.stream):
            self.seed.zero_()
            self._to_pitched(grad_u.to(e.device, torch.float64), self.seed)
            self.grad_f.zero_()
            for t in self.lam:
                t.zero_()
        if self._bwd_graph is None:
            self._bwd_graph = self._capture(self._issue_backward)
        nt.check(e.lib.d2q9_graph_launch(self._bwd_graph, e._stream_ptr),
                 "graph launch")
        with torch.cuda.stream(e.stream):
            gf = self._from_pitched(self.grad_f, 0, 2)
            gu = self._from_pitched(self.grad_init, 1, 2)
        e.stream.synchronize()
        return gf, gu

    def close(self):
        for g in (self._fwd_graph, self._bwd_graph):
            if g is not None:
                self.eng.lib.d2q9_graph_destroy(g)
        self._fwd_graph = self._bwd_graph = None


class LBMFlow(torch.autograd.Function):
    """u_N = LBMFlow.apply(force, u0, rollout): differentiable in force and
    u0 through the rollout's adjoint."""

    @staticmethod
    def forward(ctx, force, u0, rollout):
        ctx.rollout = rollout
        return rollout.forward(force.detach(), u0.detach())

    @staticmethod
    def backward(ctx, grad_u):
        gf, gu = ctx.rollout.backward(grad_u.detach().contiguous())
        return gf, gu, None
