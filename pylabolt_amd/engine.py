"""The single-lattice D2Q9 engine: device buffers, cell tables, step control.

HBM layout (one lattice):

* populations: two buffers (ping-pong) of the storage type, blocked
  (array of structures of arrays): padded cell ``(i, j)`` (``0 <= i <=
  nx+1``, ``0 <= j <= ny+1``, ghost ring included), direction ``k`` sits
  at ``((i * nbj + jj // B) * 9 + k) * B + jj % B`` with ``jj = j + B - 1``
  and ``B = 64 * cpl`` cells per block (``cpl`` 1 for FP64, 2 otherwise),
  so interior cell ``j = 1`` starts a block and one wave reads or writes
  one contiguous block of one direction;
* side planes (flags, boundary slots, moments, force field): padded cell
  ``(i, j)`` at ``origin + i * pitch + j`` with ``origin = 63`` and
  ``pitch`` a multiple of 64;
* moments: two buffers of three planes ``[rho | ux | uy]`` in the compute
  type (double for FP64 storage, float otherwise);
* one flag byte per padded cell, the boundary tables (a slot per boundary
  node, one rule per direction), and the tile tables of ``CaseTables``:
  tile list, masked-cell bits of the bulk tiles, straight-wall masks and
  the fix-up cell list.

Step bookkeeping.  The reference step n computes moments and collision from
``pop_new[n-1]`` into ``pop[n]``, then streams and applies boundaries into
``pop_new[n]`` (pylabolt/solvers/fluidLB.py:206-253).  The fused kernel
streams ``pop[n]`` and collides the result in one sweep, so after step n
the engine holds ``pop[n]`` and defers the stream; ``materialize()`` runs
the stream-only sweep when ``pop_new[n]`` is wanted.  The values are the
reference's either way.
"""
import ctypes
import math
import warnings
from types import SimpleNamespace

import numpy as np
import torch

from pylabolt_amd import _native as nt

ORIGIN = 63
ALIGN = 64

_TORCH_DTYPES = {
    nt.STORE_F64: (torch.float64, torch.float64),
    nt.STORE_F32: (torch.float32, torch.float32),
    nt.STORE_F16DEV: (torch.float16, torch.float32),
    nt.STORE_BF16DEV: (torch.bfloat16, torch.float32),
}

# D2Q9 directions (reference base/lattice.py:50-59)
CX = np.array([0, 1, 0, -1, 0, 1, -1, -1, 1])
CY = np.array([0, 0, 1, 0, -1, 1, 1, -1, -1])
_KIND = {"bounce_back": nt.RULE_BB, "fixed_velocity": nt.RULE_VEL,
         "fixed_pressure": nt.RULE_PRES, "zero_gradient": nt.RULE_ZG}


def geometry(nx, ny):
    """(pitch, plane) of the pitched layout for an nx x ny interior."""
    pitch = int(math.ceil((ny + 2) / ALIGN) * ALIGN)
    plane = (nx + 2) * pitch + ALIGN
    return pitch, plane


# default threshold of the two-step sweep, on both population buffers
# (measured on MRT channels: even from the 256 MiB last-level cache, 90 MiB
# and more ran 6-19 % faster with it; 54 MiB broke even; 18-36 MiB ran up to
# 1.8x slower, the rim's extra launches costing more than the traffic saved)
TWO_STEP_MIN_BYTES = 64 << 20

CPL = {nt.STORE_F64: 1, nt.STORE_F32: 2, nt.STORE_F16DEV: 2,
       nt.STORE_BF16DEV: 2}


def pop_geometry(nx, ny, cpl):
    """(B, nbj, pop_len) of the AoSoA population layout."""
    B = 64 * cpl
    nbj = (ny + B) // B + 1
    return B, nbj, (nx + 2) * nbj * 9 * B


def pop_aos(buf, nx, ny, B, nbj):
    """AoSoA buffer -> reference AoS view [(nx+2)*(ny+2), 9] (a copy)."""
    t = buf.reshape(nx + 2, nbj, 9, B).permute(0, 2, 1, 3)
    t = t.reshape(nx + 2, 9, nbj * B)[:, :, B - 1:B - 1 + ny + 2]
    return t.permute(0, 2, 1).reshape(-1, 9)


def padded_view(flat, q, nx, ny, pitch, plane):
    """View of plane ``q`` of ``flat`` as the padded (nx+2, ny+2) grid."""
    base = q * plane + ORIGIN
    return flat[base:base + (nx + 2) * pitch].reshape(nx + 2, pitch)[:, :ny + 2]


def in_bulk_cells(tiles, tile_i, tile_j, nx, ny):
    """Interior cells (nx, ny) lying in a bulk tile."""
    return np.repeat(np.repeat(tiles, tile_i, axis=0), tile_j,
                     axis=1)[:nx, :ny]


class CaseTables:
    """Host-built per-cell flags, boundary rules and tile classes.

    Built from the reference-shaped case arrays: ``solid`` over the padded
    grid, the ordered boundary elements and the periodic flags.
    """

    def __init__(self, nx, ny, solid_pad, elements, x_periodic, y_periodic,
                 tile_j=256, tile_i=1, solid_tiles="sweep"):
        self.nx, self.ny = nx, ny
        self.x_periodic, self.y_periodic = bool(x_periodic), bool(y_periodic)
        nxp, nyp = nx + 2, ny + 2
        solid = np.asarray(solid_pad, dtype=bool).reshape(nxp, nyp)
        ghost = np.zeros((nxp, nyp), dtype=bool)
        ghost[0, :] = ghost[-1, :] = ghost[:, 0] = ghost[:, -1] = True
        fluid = ~solid & ~ghost
        near = np.zeros_like(fluid)
        for k in range(1, 9):
            sh = np.zeros_like(solid)
            # neighbour pulled from: (i - cx, j - cy)
            src = np.roll(np.roll(solid, CX[k], axis=0), CY[k], axis=1)
            sh[1:-1, 1:-1] = src[1:-1, 1:-1]
            near |= sh
        near &= fluid
        # boundary rules, last element wins per (node, direction)
        rules = {}
        self.elems = []
        store = near.copy()
        for e in elements:
            if e["type"] == "periodic":
                continue
            kind = _KIND[e["type"]]
            eidx = len(self.elems)
            self.elems.append((kind, e))
            nrm = [int(v) for v in e["normal"]]
            for node in np.asarray(e["nodes"], dtype=np.int64):
                i, j = divmod(int(node), nyp)
                if not fluid[i, j]:
                    continue
                r = rules.setdefault((i, j), [0] * 9)
                for v in e["inv"]:
                    r[int(v)] = kind | (eidx << 4)
                if kind in (nt.RULE_VEL, nt.RULE_PRES):
                    store[i, j] = True
                if kind == nt.RULE_PRES:
                    store[i + nrm[0], j + nrm[1]] = True
                if kind == nt.RULE_ZG and near[i + nrm[0], j + nrm[1]]:
                    store[i + nrm[0], j + nrm[1]] = True
        if len(self.elems) >= 4096:
            raise ValueError("at most 4095 boundary elements are supported")
        bc = np.zeros_like(fluid)
        for (i, j) in rules:
            bc[i, j] = True
        store &= fluid
        flags = (fluid * nt.CELL_FLUID + solid * nt.CELL_SOLID +
                 near * nt.CELL_NEAR_SOLID + bc * nt.CELL_BC +
                 store * nt.CELL_STORE_MOM).astype(np.uint8)
        self.flags_pad = flags
        self.fluid_pad = fluid
        keys = sorted(rules.keys())
        self.bc_nodes = np.zeros((max(1, len(keys)), 10), dtype=np.uint16)
        self.bc_slot_pad = np.zeros((nxp, nyp), dtype=np.int32)
        for n, key in enumerate(keys):
            self.bc_nodes[n, :9] = rules[key]
            self.bc_slot_pad[key] = n
        self.n_bc_nodes = len(keys)
        # straight-wall cells whose only rules are one wall's: the bulk
        # sweep can treat them itself (sides of the device's wall masks:
        # 0/1 zero-gradient from j + 1 / j - 1, 2..5 bounce-back on the
        # walls of normal (0,1), (0,-1), (1,0), (-1,0))
        self.wall_side_pad = np.full((nxp, nyp), -1, dtype=np.int8)
        zg_sides = {((0, 1), (2, 5, 6)): 0, ((0, -1), (4, 7, 8)): 1}
        bb_sides = {((0, 1), (2, 5, 6)): 2, ((0, -1), (4, 7, 8)): 3,
                    ((1, 0), (1, 5, 8)): 4, ((-1, 0), (3, 6, 7)): 5}
        for (i, j), r in rules.items():
            dirs = tuple(k for k in range(9) if r[k] & 15 != nt.RULE_STREAM)
            kinds = {r[k] & 15 for k in dirs}
            normals = {tuple(int(v) for v in self.elems[r[k] >> 4][1]["normal"])
                       for k in dirs}
            if len(kinds) != 1 or len(normals) != 1:
                continue
            key = (normals.pop(), dirs)
            kind = kinds.pop()
            if kind == nt.RULE_ZG and key in zg_sides:
                self.wall_side_pad[i, j] = zg_sides[key]
            elif kind == nt.RULE_BB and key in bb_sides:
                self.wall_side_pad[i, j] = bb_sides[key]
        self.set_tile(tile_j, tile_i, solid_tiles)

    def set_tile(self, tile_j, tile_i, solid_tiles="sweep"):
        """Classify tiles of tile_j x tile_i interior cells.

        Bulk tiles (class 0) hold no solid cell: the bulk sweep treats all
        their cells as plain interior fluid and stores only the cells their
        mask leaves set.  Masked out are the special cells -- solid
        neighbour, boundary rule, stored moment, periodic edge -- which go
        to the fix-up list, and the cells past the lattice's last column of
        a partial tile, which are never stored (the ghost column keeps its
        initial zero; rows past the last row are not swept).  General tiles
        (class 1) hold a solid cell.

        ``solid_tiles``: "sweep" lists the general tiles for the general
        sweep; "cells" lists none and puts their fluid cells in the fix-up
        list instead (solid cells keep the encoded zero written at
        initialisation), so a step needs no general sweep; "auto" takes
        "cells" when at most 1/32 of the tiles are general.
        """
        if tile_j % 64 or not 1 <= tile_i <= 4:
            raise ValueError("tile_j must be 64 * cpl and tile_i 1..4")
        self.tile_j, self.tile_i = int(tile_j), int(tile_i)
        nx, ny = self.nx, self.ny
        inner = self.flags_pad[1:-1, 1:-1]
        fluid = (inner & nt.CELL_FLUID) != 0
        special = inner != nt.CELL_FLUID
        if self.x_periodic:
            special[0, :] = True
            special[-1, :] = True
        if self.y_periodic:
            special[:, 0] = True
            special[:, -1] = True
        special &= fluid
        # straight-wall cells the bulk sweep handles itself: plain boundary
        # cells (no solid neighbour, no stored moment, no periodic edge);
        # for zero-gradient, the source cell must be a plain pull of the
        # same block and not itself such a cell
        side = self.wall_side_pad[1:-1, 1:-1]

        def inline_walls(edge):
            plain_bc = (inner == (nt.CELL_FLUID | nt.CELL_BC)) & ~edge
            cand = plain_bc & (side >= 0) & (side <= 1)
            zg = plain_bc & (side >= 2)
            for s_, step in ((0, 1), (1, -1)):
                ci, cj = np.nonzero(cand & (side == s_))
                sj = cj + step
                ok = (sj >= 0) & (sj < ny)
                ci, cj, sj = ci[ok], cj[ok], sj[ok]
                src_flags = inner[ci, sj]
                ok = (((src_flags & nt.CELL_FLUID) != 0) &
                      ((src_flags & nt.CELL_NEAR_SOLID) == 0) &
                      ~edge[ci, sj] & ~cand[ci, sj] &
                      (cj // self.tile_j == sj // self.tile_j))
                zg[ci[ok], cj[ok]] = True
            return zg

        edge_y = np.zeros_like(fluid)
        if self.y_periodic:
            edge_y[:, 0] = edge_y[:, -1] = True
        edge = edge_y.copy()
        if self.x_periodic:
            edge[0, :] = edge[-1, :] = True
        zg = inline_walls(edge)
        self.wall_inline = zg
        # the two-step sweep reads the rows past a periodic lattice's ends
        # from their images, so there the wall cells of the first and last
        # rows are straight-wall cells too
        self.wall_inline_wrap = inline_walls(edge_y) if self.x_periodic else zg
        special &= ~zg
        ntj = -(-ny // self.tile_j)
        nti = -(-nx // self.tile_i)
        # cells past the lattice count as fluid here: they do not stop a
        # tile from being swept as bulk
        padded = np.ones((nti * self.tile_i, ntj * self.tile_j), dtype=bool)
        padded[:nx, :ny] = fluid
        tiles = padded.reshape(nti, self.tile_i, ntj, self.tile_j).all(axis=(1, 3))
        self.tile_class = np.where(tiles, 0, 1).astype(np.uint8).ravel()
        self.ntj, self.nti = ntj, nti
        ids = np.arange(self.tile_class.size, dtype=np.int32)
        bulk_ids = ids[self.tile_class == 0]
        general_ids = ids[self.tile_class != 0]
        if solid_tiles not in ("sweep", "cells", "auto"):
            raise ValueError("solid_tiles must be 'sweep', 'cells' or 'auto'")
        self.solid_tiles_req = solid_tiles
        mode = solid_tiles
        if mode == "auto":
            mode = "cells" if 32 * general_ids.size <= ids.size else "sweep"
        self.solid_tiles = mode
        if mode == "sweep":
            self.tile_list = np.concatenate([bulk_ids, general_ids]).astype(np.int32)
            self.n_general = int(general_ids.size)
        else:
            self.tile_list = bulk_ids.astype(np.int32)
            self.n_general = 0
        self.n_bulk = int(bulk_ids.size)
        # fix-up cells: special cells lying in bulk tiles ("cells": and every
        # fluid cell of a tile holding a solid)
        in_bulk = np.repeat(np.repeat(tiles, self.tile_i, axis=0),
                            self.tile_j, axis=1)[:nx, :ny]
        # masked bits per (bulk tile, row, 64-cell word): special cells
        # and cells past the lattice
        spad = np.ones((nti * self.tile_i, ntj * self.tile_j), dtype=bool)
        spad[:nx, :ny] = special
        words = self.tile_j // 64
        bits = spad.reshape(nti, self.tile_i, ntj, words, 64)
        weights = (np.uint64(1) << np.arange(64, dtype=np.uint64))
        m = (bits.astype(np.uint64) * weights).sum(axis=-1, dtype=np.uint64)
        m = m.transpose(0, 2, 1, 3).reshape(nti * ntj, self.tile_i, words)
        self.bulk_masks = np.ascontiguousarray(m[bulk_ids]).reshape(-1)
        if self.bulk_masks.size == 0:
            self.bulk_masks = np.zeros(1, dtype=np.uint64)
        # wall bits per (bulk tile, row, side, word)
        zm = []
        for s_ in range(6):
            zpad = np.zeros((nti * self.tile_i, ntj * self.tile_j), dtype=bool)
            zpad[:nx, :ny] = zg & (side == s_)
            zb = zpad.reshape(nti, self.tile_i, ntj, words, 64)
            w_ = (zb.astype(np.uint64) * weights).sum(axis=-1, dtype=np.uint64)
            zm.append(w_.transpose(0, 2, 1, 3).reshape(nti * ntj, self.tile_i,
                                                       words))
        zmask = np.stack(zm, axis=2)[bulk_ids]   # [bulk, row, side, word]
        self.n_wall = int((zg & in_bulk_cells(tiles, self.tile_i, self.tile_j,
                                              nx, ny)).sum())
        # compact: only the bulk tiles holding wall cells get a slot
        has = zmask.reshape(zmask.shape[0], -1).any(axis=1)
        self.wall_index = np.where(has, np.cumsum(has) - 1, -1).astype(np.int32)
        self.wall_masks = (np.ascontiguousarray(zmask[has]).reshape(-1)
                           if self.n_wall else None)
        fixed = special & in_bulk
        if mode == "cells":
            fixed |= fluid & ~in_bulk
        ii, jj = np.nonzero(fixed)
        pitch = int(math.ceil((ny + 2) / ALIGN) * ALIGN)
        self.fix_cells = (ORIGIN + (ii + 1) * pitch + (jj + 1)).astype(np.int32)
        if self.fix_cells.size == 0:
            self.fix_cells = np.zeros(1, dtype=np.int32)
            self.n_fix = 0
        else:
            self.n_fix = int(self.fix_cells.size)
        return self.tile_class


    def deep_split(self, rows):
        """Deep tiles of the two-step sweep (d2q9_fused_step2) and the rim
        tables around them.

        A deep tile is ``rows`` x one block (``tile_j`` cells) of interior
        cells (block-aligned, rows a multiple of the tile rows) whose cells
        and neighbours one cell deep are all plain fluid -- no boundary
        rule, solid neighbour, stored moment or periodic edge, and inside
        the lattice, where the rows past the ends of a lattice periodic in
        x with at least 2 ``rows`` rows are their periodic images (and its
        last row of tiles may be moved up, see below) -- except that the
        first or last column of the lattice
        may be a straight zero-gradient or bounce-back wall swept in the
        tile (wall rows of sides 0-3, one kind along the tile's rows; then
        the ghost column beside it is the one cell of the grown region
        outside the lattice; ``deep_walls`` gives the kinds).  The rim is every
        other cell, swept by whole tiles of the current tile shape (a tile
        is either inside one deep tile or in the rim).  Rim B lists the rim
        tiles with their fix-up cells; rim A lists the same tiles and adds
        to the fix-up list every deep cell a rim cell pulls from, so that
        its first step leaves in the scratch lattice all that the second
        reads.  Returns a namespace (deep_ids, deep_walls, n_deep,
        tile_list, n_bulk, n_general, bulk_masks, wall_index, fix_a,
        fix_b)."""
        B = self.tile_j
        if rows % self.tile_i:
            raise ValueError("deep tile rows must be a multiple of the tile "
                             "rows")
        nx, ny = self.nx, self.ny
        inner = self.flags_pad[1:-1, 1:-1]
        plain = inner == nt.CELL_FLUID
        # In a lattice periodic in x with at least two tile heights of rows,
        # the deep kernel reads the rows past the first and last from their
        # periodic images, so the edge rows are plain; and when the rows are
        # a multiple of the tile rows, the last row of deep tiles is moved to
        # end at the last row (overlapping the one before it), as the kernel
        # does for a tile that would pass it.  A channel periodic in x with
        # straight walls at y is then deep throughout: no rim.
        wrap = bool(self.x_periodic) and nx >= 2 * rows
        if self.x_periodic and not wrap:
            plain[0, :] = plain[-1, :] = False
        if self.y_periodic:
            plain[:, 0] = plain[:, -1] = False
        ext = (lambda a: np.concatenate([a[-1:], a, a[:1]], axis=0)) if wrap \
            else (lambda a: a)
        off = 1 if wrap else 0          # row r (0-based) is row r + off of ext(...)
        plain_x = ext(plain)
        sat = np.zeros((plain_x.shape[0] + 1, ny + 1), dtype=np.int64)
        sat[1:, 1:] = (~plain_x).astype(np.int64).cumsum(0).cumsum(1)
        ntj = -(-ny // B)
        nti2 = -(-nx // rows)
        first = np.arange(nti2) * rows  # first row of each row of tiles
        if wrap and nx % self.tile_i == 0:
            first = np.minimum(first, nx - rows)
        ti = first[:, None]
        tj = np.arange(ntj)[None, :]
        # grown region (0-based, inclusive): rows first - 1 .. first + rows,
        # cells tj*B - 1 .. tj*B + B
        r0, r1 = ti - 1, ti + rows
        c0, c1 = tj * B - 1, tj * B + B
        # a tile at the low (high) wall: its grown region reaches the ghost
        # column, and the wall column is checked apart from the plain cells
        at_lo = (c0 == -1) & np.ones_like(r0, dtype=bool)
        at_hi = (c1 == ny) & np.ones_like(r0, dtype=bool)
        ca = np.where(at_lo, 1, c0)
        cb = np.where(at_hi, ny - 2, c1)
        rows_ok = (r1 - 1 <= nx - 1) if wrap else ((r0 >= 0) & (r1 <= nx - 1))
        ok = rows_ok & (ca >= 0) & (cb <= ny - 1) & (ca <= cb)
        nxe = plain_x.shape[0]
        r0c, r1c = np.clip(r0 + off, 0, nxe - 1), np.clip(r1 + off, 0, nxe - 1)
        cac, cbc = np.clip(ca, 0, ny - 1), np.clip(cb, 0, ny - 1)
        bad = (sat[r1c + 1, cbc + 1] - sat[r0c, cbc + 1] - sat[r1c + 1, cac]
               + sat[r0c, cac])
        ok &= bad == 0
        # wall kinds per row: 1 zero-gradient, 2 bounce-back, 0 neither (a
        # wall cell the bulk sweep does inline, sides 0/2 at j = 1, 1/3 at
        # j = ny)
        side = self.wall_side_pad[1:-1, 1:-1]
        inl = self.wall_inline_wrap if wrap else self.wall_inline
        lo_row = np.where(inl[:, 0] & (side[:, 0] == 0), 1,
                          np.where(inl[:, 0] & (side[:, 0] == 2), 2, 0))
        hi_row = np.where(inl[:, -1] & (side[:, -1] == 1), 1,
                          np.where(inl[:, -1] & (side[:, -1] == 3), 2, 0))
        n_rows = r1c - r0c + 1

        def kind_over_rows(per_row):
            kind = np.zeros_like(r0)
            per_row = ext(per_row)
            for k in (1, 2):
                cum = np.concatenate([[0], np.cumsum(per_row == k)])
                full = (cum[r1c + 1] - cum[r0c]) == n_rows
                kind = np.where(full, k, kind)
            return kind

        lo_kind = np.where(at_lo, kind_over_rows(lo_row), 0)
        hi_kind = np.where(at_hi, kind_over_rows(hi_row), 0)
        ok &= ~at_lo | (lo_kind > 0)
        ok &= ~at_hi | (hi_kind > 0)
        tix = np.arange(nti2)[:, None]
        deep_ids = (tix * ntj + tj)[ok].astype(np.int32)
        deep_walls = (lo_kind | (hi_kind << 2))[ok].astype(np.uint8)
        covered = np.zeros((nx, ny), dtype=bool)
        for a in range(nti2):
            covered[first[a]:first[a] + rows] |= np.repeat(ok[a], B)[:ny]
        # rim tiles of the current tile shape
        nti = self.nti
        cpad = np.zeros((nti * self.tile_i, ntj * B), dtype=bool)
        cpad[:nx, :ny] = covered
        in_deep = cpad.reshape(nti, self.tile_i, ntj, B).all(axis=(1, 3))
        in_deep = in_deep.ravel()
        tl = np.asarray(self.tile_list)
        bulk_pos = np.nonzero(~in_deep[tl[:self.n_bulk]])[0]
        gen = tl[self.n_bulk:self.n_bulk + self.n_general]
        gen = gen[~in_deep[gen]]
        words = self.tile_j // 64
        if self.n_bulk:
            masks = self.bulk_masks.reshape(self.n_bulk, self.tile_i, words)
            bulk_masks = np.ascontiguousarray(masks[bulk_pos]).reshape(-1)
        else:
            bulk_masks = np.zeros(0, dtype=np.uint64)
        if bulk_masks.size == 0:
            bulk_masks = np.zeros(1, dtype=np.uint64)
        wall_index = (self.wall_index[bulk_pos].astype(np.int32)
                      if self.wall_masks is not None else None)
        if wall_index is not None and wall_index.size == 0:
            # no rim bulk tile: a placeholder, so that the array is not null
            wall_index = np.full(1, -1, dtype=np.int32)
        pitch = int(math.ceil((ny + 2) / ALIGN) * ALIGN)

        def cells_of(c):
            c = np.asarray(c, dtype=np.int64) - ORIGIN
            return c // pitch - 1, c % pitch - 1

        fi, fj = cells_of(self.fix_cells[:self.n_fix])
        keep = ~covered[fi, fj] if self.n_fix else np.zeros(0, dtype=bool)
        fix_b = np.asarray(self.fix_cells[:self.n_fix])[keep].astype(np.int64)
        # deep cells a rim fluid cell pulls from (periodic wrap included)
        rim = ~covered & ((inner & nt.CELL_FLUID) != 0)
        near = np.zeros_like(rim)
        for di in (-1, 0, 1):
            for dj in (-1, 0, 1):
                sh = np.roll(rim, (di, dj), axis=(0, 1))
                if not self.x_periodic:
                    if di == 1:
                        sh[0, :] = False
                    elif di == -1:
                        sh[-1, :] = False
                if not self.y_periodic:
                    if dj == 1:
                        sh[:, 0] = False
                    elif dj == -1:
                        sh[:, -1] = False
                near |= sh
        ri, rj = np.nonzero(near & covered)
        ring = ORIGIN + (ri + 1).astype(np.int64) * pitch + (rj + 1)
        fix_a = np.union1d(fix_b, ring)

        def as_list(c):
            c = np.asarray(c, dtype=np.int32)
            return (c, int(c.size)) if c.size else (np.zeros(1, np.int32), 0)

        tile_list = np.concatenate([tl[:self.n_bulk][bulk_pos],
                                    gen]).astype(np.int32)
        if tile_list.size == 0:
            tile_list = np.zeros(1, dtype=np.int32)
        fa, nfa = as_list(fix_a)
        fb, nfb = as_list(fix_b)
        return SimpleNamespace(
            deep_ids=deep_ids if deep_ids.size else np.zeros(1, np.int32),
            deep_walls=(deep_walls if deep_walls.size
                        else np.zeros(1, np.uint8)),
            n_deep=int(deep_ids.size), rows=int(rows), covered=covered,
            tile_list=tile_list, n_bulk=int(bulk_pos.size),
            n_general=int(gen.size), bulk_masks=bulk_masks,
            wall_index=wall_index, fix_a=fa, n_fix_a=nfa, fix_b=fb,
            n_fix_b=nfb)


def force_tables(nx, ny, groups):
    """Device-ready node and group records for d2q9_momentum_exchange.

    ``groups``: list of dicts with ``nodes`` (reference padded indices
    ``i * (ny + 2) + j``), ``masks`` (direction bits per node), ``ref``
    (x, y) and ``torque`` (bool).  Each group's nodes fill whole blocks of
    256 records; padding records have mask 0."""
    pitch, _ = geometry(nx, ny)
    nyp = ny + 2
    recs, grecs = [], np.zeros(len(groups), dtype=nt.FORCE_GROUP_DTYPE)
    block = 0
    for g, spec in enumerate(groups):
        nodes = np.asarray(spec["nodes"], dtype=np.int64)
        masks = np.asarray(spec["masks"], dtype=np.uint32)
        n = int(nodes.size)
        nb = -(-n // nt.FORCE_BLOCK)
        r = np.zeros(nb * nt.FORCE_BLOCK, dtype=nt.FORCE_NODE_DTYPE)
        i, j = nodes // nyp, nodes % nyp
        r["cell"][:n] = ORIGIN + i * pitch + j
        r["group"][:] = g
        r["mask"][:n] = masks
        recs.append(r)
        grecs[g]["ref"] = spec.get("ref", (0.0, 0.0))
        grecs[g]["torque"] = int(bool(spec.get("torque", False)))
        grecs[g]["first_block"] = block
        grecs[g]["n_blocks"] = nb
        block += nb
    nodes = (np.concatenate(recs) if recs else
             np.zeros(0, dtype=nt.FORCE_NODE_DTYPE))
    return nodes, grecs


def elements_from_boundary(boundary):
    """Ordered element dicts from a ``Boundary`` (reference attribute
    names, base/boundary.py:7-84)."""
    out = []
    for e in boundary.boundary_elements:
        out.append(dict(
            type=e.type_fluid, nodes=np.asarray(e.boundary_nodes, np.int64),
            out=[int(v) for v in e.out_list], inv=[int(v) for v in e.inv_list],
            vector=[float(v) for v in e.vector_fluid],
            scalar=float(e.scalar_fluid),
            normal=[int(round(float(v))) for v in e.surface_normals]))
    return out


class D2Q9Engine:
    """Device state and kernel calls of one lattice on one GPU.

    Parameters mirror the reference's case description: padded ``solid``
    mask, initial ``density``/``velocity`` (reference layouts), ordered
    boundary ``elements``, periodic flags, collision ``model`` ("BGK" or
    "MRT"), ``forcing`` (None, "guo_linear", "guo_second_order"),
    ``omega``, ``gravity``; plus the storage form ("f64", "f32", "f16dev",
    "bf16dev") and an optional per-cell ``force_field`` (padded
    ``[size, 2]``) that replaces ``rho * gravity``.  ``solid_tiles`` picks
    how tiles holding solids are swept (see ``CaseTables.set_tile``);
    ``two_step`` advances pairs of steps with one sweep over the deep
    interior (d2q9_fused_step2) when deep tiles fit; by default for FP64
    and FP32 storage when the two population buffers exceed
    ``TWO_STEP_MIN_BYTES`` (the results are bitwise those of single
    steps), for 16-bit storage only when asked (its intermediate step stays
    in FP32, which changes the rounding).
    """

    def __init__(self, nx, ny, solid_pad, density, velocity, elements,
                 x_periodic, y_periodic, model, forcing, omega, gravity,
                 inv_cs_2, inv_cs_4, weights, float_min, storage="f64",
                 force_field=None, tile=(64, 4), rates=None, dev_scale=None,
                 device=None, graph_steps=16, stream=None, solid_tiles=None,
                 two_step=None):
        if force_field is not None and forcing is None:
            # a body force acts through the Guo source term
            raise ValueError("a force field needs a Guo forcing model")
        self.lib = nt.lib()
        self.device = (torch.device(device) if device is not None else
                       torch.device("cuda", torch.cuda.current_device()))
        self.nx, self.ny = int(nx), int(ny)
        self.pitch, self.plane = geometry(self.nx, self.ny)
        self.storage_name = storage
        self.storage = nt.STORAGE_CODES[storage]
        self.raw_dtype, self.ctype = _TORCH_DTYPES[self.storage]
        self.model = model
        self.forcing = forcing
        self.weights = np.asarray(weights, dtype=np.float64)
        self.cpl = CPL[self.storage]
        self.B, self.nbj, self.pop_len = pop_geometry(self.nx, self.ny,
                                                      self.cpl)
        self.tile_i = int(tile[1]) if tile is not None else 4
        if solid_tiles is None:
            # 16-bit steps run as one kernel when the tiles holding solids
            # are few (their fluid cells go through the fix-up path)
            solid_tiles = ("auto" if self.storage in (nt.STORE_F16DEV,
                                                      nt.STORE_BF16DEV)
                           else "sweep")
        self.tables = CaseTables(self.nx, self.ny, solid_pad, elements,
                                 x_periodic, y_periodic, self.B, self.tile_i,
                                 solid_tiles)
        self.stream = (stream if stream is not None else
                       torch.cuda.Stream(device=self.device))
        self._stream_ptr = ctypes.c_void_p(self.stream.cuda_stream)
        dev = self.device
        P = self.plane
        with torch.cuda.stream(self.stream):
            self.f = [torch.empty(self.pop_len, dtype=self.raw_dtype,
                                  device=dev) for _ in range(2)]
            self.mom = [torch.zeros(3 * P, dtype=self.ctype, device=dev)
                        for _ in range(2)]
            self.flags = self._upload_plane(self.tables.flags_pad, torch.uint8)
            self.bc_slot = self._upload_plane(self.tables.bc_slot_pad,
                                              torch.int32)
            self.bc_nodes = torch.from_numpy(
                self.tables.bc_nodes.view(np.int16).copy()).to(dev)
            elems = (nt.BcElem * max(1, len(self.tables.elems)))()
            for n, (kind, e) in enumerate(self.tables.elems):
                elems[n].vec[0], elems[n].vec[1] = e["vector"]
                elems[n].scalar = e["scalar"]
                elems[n].normal[0], elems[n].normal[1] = e["normal"]
                elems[n].type = kind
            raw = bytes(elems)
            self.bc_elems = torch.frombuffer(bytearray(raw),
                                             dtype=torch.uint8).to(dev)
            self.force_field = None
            if force_field is not None:
                self.force_field = self._pitched_vector(force_field)
            self.work = torch.zeros(6 * 2048, dtype=torch.float64, device=dev)
            self.red_out = torch.zeros(8, dtype=torch.float64, device=dev)
        L = nt.Lattice()
        L.nx, L.ny, L.pitch, L.origin, L.plane = (self.nx, self.ny, self.pitch,
                                                 ORIGIN, P)
        L.cpl, L.nbj, L.pop_len = self.cpl, self.nbj, self.pop_len
        L.storage = self.storage
        L.collision = nt.MODEL_CODES[model]
        L.forcing = nt.FORCING_CODES[forcing]
        L.force_source = nt.FSRC_FIELD if force_field is not None else nt.FSRC_GRAVITY
        L.x_periodic, L.y_periodic = int(x_periodic), int(y_periodic)
        L.tile_j, L.tile_i = self.tables.tile_j, self.tables.tile_i
        L.omega = float(omega)
        L.inv_cs_2, L.inv_cs_4, L.float_min = (float(inv_cs_2), float(inv_cs_4),
                                               float(float_min))
        for k in range(9):
            L.w[k] = float(self.weights[k])
        L.gravity[0], L.gravity[1] = float(gravity[0]), float(gravity[1])
        if rates is None:
            rates = [1, 1, 1, 1, 1, 1, 1, omega, omega]
        if model == "MRT" and any(float(r) != 1.0 for r in rates[:7]):
            # the kernels evaluate the reference's S, whose rows 0..6 relax
            # at rate 1 (base/collision_operator.py:158-162)
            raise ValueError("MRT rates of rows 0..6 must be 1; only the "
                             "stress rows 7 and 8 take other rates")
        for k in range(9):
            L.rates[k] = float(rates[k])
        if dev_scale is None:
            # FP16 keeps (f - w) / w * 1024 clear of its subnormal range;
            # the scale is the power of two at or below 1024 / w, so that
            # decoding (times 1 / scale) undoes it exactly: a scale whose
            # inverse rounds would bias every decoded value by the same
            # relative amount, step after step
            if self.storage == nt.STORE_F16DEV:
                dev_scale = [2.0 ** math.floor(math.log2(1024.0 / float(w)))
                             for w in self.weights]
            else:
                dev_scale = [1.0] * 9
        for k in range(9):
            L.dev_scale[k] = float(dev_scale[k])
        cf = L.cf
        for k in range(9):
            cf.w[k] = L.w[k]
            cf.rates[k] = L.rates[k]
            cf.scale[k] = L.dev_scale[k]
            cf.inv_scale[k] = 1.0 / L.dev_scale[k]
        cf.omega, cf.inv_cs_2, cf.inv_cs_4 = L.omega, L.inv_cs_2, L.inv_cs_4
        cf.float_min = L.float_min
        cf.gravity[0], cf.gravity[1] = L.gravity[0], L.gravity[1]
        L.flags = self.flags.data_ptr()
        if self.tables.n_bc_nodes:
            L.bc_slot = self.bc_slot.data_ptr()
            L.bc_nodes = self.bc_nodes.data_ptr()
        L.bc_elems = self.bc_elems.data_ptr()
        L.force_field = (self.force_field.data_ptr()
                         if self.force_field is not None else None)
        self.L = L
        self._Lp = ctypes.byref(self.L)
        self._upload_tiles()
        self.graph_steps = max(2, int(graph_steps) // 2 * 2)
        self._graphs = {}
        # pairs of steps with one sweep over the deep interior, by default
        # for lattices whose two population buffers exceed
        # TWO_STEP_MIN_BYTES
        if two_step is None:
            # FP64 and FP32 give the same bits either way; in 16-bit
            # storage the two-step sweep keeps the intermediate step in FP32
            # and so changes the numerics, which only an explicit request
            # may do
            two_step = (self.storage in (nt.STORE_F64, nt.STORE_F32) and
                        2 * self.f[0].numel() * self.f[0].element_size()
                        > TWO_STEP_MIN_BYTES)
        self._two_step_req = bool(two_step)
        self.two_step = False
        self._setup_two_step()
        self.time_step = 0
        self._init_density = np.asarray(density, dtype=np.float64)
        self._init_velocity = np.asarray(velocity, dtype=np.float64)
        self.initialize()

    def _upload_tiles(self):
        """Copy the tile list, bulk masks and fix-up cells of the current
        tile shape to the device and point the descriptor at them."""
        dev = self.device
        with torch.cuda.stream(self.stream):
            self.tile_list = torch.from_numpy(self.tables.tile_list).to(dev)
            self.fix_cells = torch.from_numpy(self.tables.fix_cells).to(dev)
            self.bulk_masks = torch.from_numpy(
                self.tables.bulk_masks.view(np.int64)).to(dev)
        L = self.L
        L.tile_j, L.tile_i = self.tables.tile_j, self.tables.tile_i
        L.tiles = self.tile_list.data_ptr()
        L.n_bulk, L.n_general = self.tables.n_bulk, self.tables.n_general
        L.fix_cells = self.fix_cells.data_ptr()
        L.n_fix = self.tables.n_fix
        L.bulk_masks = self.bulk_masks.data_ptr()
        self.wall_masks = self.wall_index = None
        if self.tables.wall_masks is not None:
            with torch.cuda.stream(self.stream):
                self.wall_masks = torch.from_numpy(
                    self.tables.wall_masks.view(np.int64)).to(self.device)
                self.wall_index = torch.from_numpy(
                    self.tables.wall_index).to(self.device)
        L.wall_masks = (self.wall_masks.data_ptr()
                        if self.wall_masks is not None else None)
        L.wall_index = (self.wall_index.data_ptr()
                        if self.wall_index is not None else None)

    def _setup_two_step(self):
        """Deep tiles and rim descriptors of the two-step sweep, and its
        scratch lattice; two-step stays off when no deep tile fits."""
        self.two_step = False
        if not self._two_step_req:
            return
        d = self.tables.deep_split(int(self.lib.d2q9_deep_rows()))
        if d.n_deep == 0:
            return
        dev = self.device
        with torch.cuda.stream(self.stream):
            self._deep = {
                "deep": torch.from_numpy(d.deep_ids).to(dev),
                "walls": torch.from_numpy(d.deep_walls).to(dev),
                "tiles": torch.from_numpy(d.tile_list).to(dev),
                "masks": torch.from_numpy(d.bulk_masks.view(np.int64)).to(dev),
                "fix_a": torch.from_numpy(d.fix_a).to(dev),
                "fix_b": torch.from_numpy(d.fix_b).to(dev),
                "wall": (torch.from_numpy(d.wall_index).to(dev)
                         if d.wall_index is not None else None),
            }
            if not hasattr(self, "f_tmp"):
                # initialize() copies f[0] into it: ghost and solid cells
                # of the scratch lattice hold the encoded zero, like those
                # of f[0] and f[1].  Without room for it the engine keeps
                # single steps.
                try:
                    self.f_tmp = torch.empty_like(self.f[0])
                    self.mom_tmp = torch.zeros(3 * self.plane,
                                               dtype=self.ctype, device=dev)
                except torch.cuda.OutOfMemoryError:
                    self.__dict__.pop("f_tmp", None)
                    self._deep = None
                    warnings.warn("no device memory for the two-step sweep's "
                                  "scratch lattice: single steps")
                    return
        t = self._deep
        rims = []
        for fix, nfix in ((t["fix_a"], d.n_fix_a), (t["fix_b"], d.n_fix_b)):
            R = nt.Lattice.from_buffer_copy(self.L)
            R.tiles = t["tiles"].data_ptr()
            R.n_bulk, R.n_general = d.n_bulk, d.n_general
            R.bulk_masks = t["masks"].data_ptr()
            R.fix_cells = fix.data_ptr()
            R.n_fix = nfix
            R.wall_index = (t["wall"].data_ptr() if t["wall"] is not None
                            else None)
            rims.append(R)
        self._rim_a, self._rim_b = rims
        self._rim_ap = ctypes.byref(self._rim_a)
        self._rim_bp = ctypes.byref(self._rim_b)
        self.n_deep = d.n_deep
        self.deep_fraction = float(d.covered.mean())
        self.two_step = True
        # a graph replays an even number of two-step sweeps, so the buffers
        # return to their roles
        self.graph_steps = max(4, self.graph_steps // 4 * 4)

    def set_tile_rows(self, tile_i):
        """Re-tile with ``tile_i`` rows per tile (workgroup 64 x tile_i);
        drops captured graphs, which hold the old tile tables."""
        tile_i = int(tile_i)
        if tile_i == self.tile_i:
            return
        self.stream.synchronize()
        self.close()
        self.tables.set_tile(self.B, tile_i, self.tables.solid_tiles_req)
        self.tile_i = tile_i
        self._upload_tiles()
        if self.two_step:
            self._setup_two_step()

    # ---- layout helpers -------------------------------------------------
    def _upload_plane(self, pad, dtype, count=1):
        host = np.zeros(count * self.plane, dtype=pad.dtype)
        padded_view(host, 0, self.nx, self.ny, self.pitch, self.plane)[:] = pad
        return torch.from_numpy(host).to(self.device).view(dtype)

    def _pitched_vector(self, arr_size2):
        arr = np.asarray(arr_size2, dtype=np.float64).reshape(
            self.nx + 2, self.ny + 2, 2)
        host = np.zeros(2 * self.plane, dtype=np.float64)
        for q in range(2):
            padded_view(host, q, self.nx, self.ny, self.pitch,
                        self.plane)[:] = arr[:, :, q]
        return torch.from_numpy(host).to(self.device, self.ctype)

    def _mom_from_host(self, density, velocity):
        host = np.zeros(3 * self.plane, dtype=np.float64)
        nxp, nyp = self.nx + 2, self.ny + 2
        padded_view(host, 0, self.nx, self.ny, self.pitch, self.plane)[:] = \
            np.asarray(density).reshape(nxp, nyp)
        vel = np.asarray(velocity).reshape(nxp, nyp, 2)
        for q in range(2):
            padded_view(host, q + 1, self.nx, self.ny, self.pitch,
                        self.plane)[:] = vel[:, :, q]
        return torch.from_numpy(host).to(self.device, self.ctype)

    def _grid(self, flat, q):
        return padded_view(flat, q, self.nx, self.ny, self.pitch, self.plane)

    # ---- lifecycle ------------------------------------------------------
    def initialize(self):
        """pop = pop_new = feq(rho0, u0) (equilibrium_kernels.py:38-78)."""
        with torch.cuda.stream(self.stream):
            m = self._mom_from_host(self._init_density, self._init_velocity)
            self.mom[0].copy_(m)
            self.mom[1].copy_(m)
        nt.check(self.lib.d2q9_init_eq(self._Lp, nt.ptr(self.f[0]),
                                       nt.ptr(self.f[1]), nt.ptr(self.mom[0]),
                                       self._stream_ptr), "d2q9_init_eq")
        if hasattr(self, "f_tmp"):
            with torch.cuda.stream(self.stream):
                self.f_tmp.copy_(self.f[0])
        # both buffers hold pop_new[0] == pop[0]
        self.i_pop = 1
        self.i_pop_new = 0
        self.i_mom = 0
        self.mom_full = True
        self.time_step = 0

    def close(self):
        for g in self._graphs.values():
            self.lib.d2q9_graph_destroy(g)
        self._graphs = {}

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- kernels --------------------------------------------------------
    def _collide(self, src, dst, write_mom):
        nt.check(self.lib.d2q9_collide(
            self._Lp, nt.ptr(self.f[src]), nt.ptr(self.f[dst]),
            nt.ptr(self.mom[self.i_mom]), nt.ptr(self.mom[1 - self.i_mom]),
            int(write_mom), self._stream_ptr), "d2q9_collide")
        self.i_mom = 1 - self.i_mom

    def _fused(self, src, dst, write_mom):
        nt.check(self.lib.d2q9_fused_step(
            self._Lp, nt.ptr(self.f[src]), nt.ptr(self.f[dst]),
            nt.ptr(self.mom[self.i_mom]), nt.ptr(self.mom[1 - self.i_mom]),
            int(write_mom), self._stream_ptr), "d2q9_fused_step")
        self.i_mom = 1 - self.i_mom

    def _fused2(self, src, dst, write_mom):
        nt.check(self.lib.d2q9_fused_step2(
            self._Lp, self._rim_ap, self._rim_bp,
            nt.ptr(self._deep["deep"]), nt.ptr(self._deep["walls"]),
            self.n_deep,
            nt.ptr(self.f[src]), nt.ptr(self.f[dst]), nt.ptr(self.f_tmp),
            nt.ptr(self.mom[self.i_mom]), nt.ptr(self.mom[1 - self.i_mom]),
            nt.ptr(self.mom_tmp), int(write_mom), self._stream_ptr),
            "d2q9_fused_step2")
        self.i_mom = 1 - self.i_mom

    def step2(self, write_moments=False):
        """Advance two reference time steps with the two-step sweep."""
        if self.i_pop_new is not None:
            self.step()
            self.step(write_moments)
            return
        src = self.i_pop
        dst = 1 - src
        self._fused2(src, dst, write_moments)
        self.i_pop = dst
        self.i_pop_new = None
        self.mom_full = bool(write_moments)
        self.time_step += 2

    def step(self, write_moments=False):
        """Advance one reference time step."""
        if self.i_pop_new is not None:
            # pop_new is in memory: moments + collision only
            src = self.i_pop_new
            dst = 1 - src
            self._collide(src, dst, write_moments)
        else:
            src = self.i_pop
            dst = 1 - src
            self._fused(src, dst, write_moments)
        self.i_pop = dst
        self.i_pop_new = None
        self.mom_full = bool(write_moments)
        self.time_step += 1

    def advance(self, n, write_moments=True, use_graph=True):
        """Advance ``n`` steps; the last one stores rho/u on every fluid
        cell when ``write_moments``.  Full ``graph_steps`` chunks of fused
        steps replay a captured HIP graph."""
        n = int(n)
        if n <= 0:
            return
        if self.i_pop_new is not None:
            self.step(write_moments and n == 1)
            n -= 1
        if self.two_step:
            # pairs of steps: graphs of graph_steps steps, the last pair
            # eager when it has to store the moments
            keep = 2 if write_moments else 0
            if use_graph:
                while n - keep >= self.graph_steps:
                    g = self._graph_for_state()
                    nt.check(self.lib.d2q9_graph_launch(g, self._stream_ptr),
                             "graph launch")
                    self.time_step += self.graph_steps
                    n -= self.graph_steps
            while n >= 2:
                self.step2(write_moments and n == 2)
                n -= 2
        # full chunks of fused steps replay a captured graph; the last step
        # stays eager when it has to store the moments
        keep = 1 if write_moments else 0
        if use_graph:
            while n - keep >= self.graph_steps:
                g = self._graph_for_state()
                nt.check(self.lib.d2q9_graph_launch(g, self._stream_ptr),
                         "graph launch")
                self.time_step += self.graph_steps
                n -= self.graph_steps
        while n > 0:
            self.step(write_moments and n == 1)
            n -= 1

    def capture_graph(self):
        """Capture (once) the graph that ``advance`` replays from the
        current buffer roles, without running it; a timed region that
        follows then replays graphs only."""
        if self.i_pop_new is None:
            full = self.mom_full
            self._graph_for_state()
            self.mom_full = full

    def _graph_for_state(self):
        key = (self.i_pop, self.i_mom)
        g = self._graphs.get(key)
        if g is None:
            saved_mom = self.i_mom
            g = self._graph_capture(self.i_pop, self.i_mom)
            self.i_mom = saved_mom
            self._graphs[key] = g
        # graph_steps is even (a multiple of 4 for two-step sweeps): the
        # buffers return to the same roles
        self.mom_full = False
        return g

    def _graph_capture(self, start, im):
        nt.check(self.lib.d2q9_graph_begin(self._stream_ptr), "graph begin")
        err = None
        src = start
        for _ in range(self.graph_steps // 2 if self.two_step
                       else self.graph_steps):
            if self.two_step:
                rc = self.lib.d2q9_fused_step2(
                    self._Lp, self._rim_ap, self._rim_bp,
                    nt.ptr(self._deep["deep"]), nt.ptr(self._deep["walls"]),
                    self.n_deep,
                    nt.ptr(self.f[src]), nt.ptr(self.f[1 - src]),
                    nt.ptr(self.f_tmp), nt.ptr(self.mom[im]),
                    nt.ptr(self.mom[1 - im]), nt.ptr(self.mom_tmp), 0,
                    self._stream_ptr)
            else:
                rc = self.lib.d2q9_fused_step(
                    self._Lp, nt.ptr(self.f[src]), nt.ptr(self.f[1 - src]),
                    nt.ptr(self.mom[im]), nt.ptr(self.mom[1 - im]), 0,
                    self._stream_ptr)
            if rc != 0 and err is None:
                err = self.lib.d2q9_last_error().decode()
            src, im = 1 - src, 1 - im
        handle = ctypes.c_void_p()
        rc_end = self.lib.d2q9_graph_end(self._stream_ptr, ctypes.byref(handle))
        if err is not None:
            raise nt.NativeError("graph capture: " + err)
        nt.check(rc_end, "graph end")
        return handle

    def materialize(self):
        """Run the deferred stream + boundary sweep: pop_new of the current
        step into the spare buffer."""
        if self.i_pop_new is not None:
            return self.i_pop_new
        dst = 1 - self.i_pop
        nt.check(self.lib.d2q9_stream_bc(
            self._Lp, nt.ptr(self.f[self.i_pop]), nt.ptr(self.f[dst]),
            nt.ptr(self.mom[self.i_mom]), self._stream_ptr), "d2q9_stream_bc")
        self.i_pop_new = dst
        return dst

    def synchronize(self):
        self.stream.synchronize()

    # ---- reductions -----------------------------------------------------
    def mass(self, which="pop"):
        """FP64 sum of the populations over fluid cells."""
        idx = self.i_pop if which == "pop" else self.materialize()
        nt.check(self.lib.d2q9_mass_fp64(
            self._Lp, nt.ptr(self.f[idx]), nt.ptr(self.work),
            nt.ptr(self.red_out), self._stream_ptr), "d2q9_mass_fp64")
        self.synchronize()
        return float(self.red_out[0].item())

    def set_force_groups(self, groups):
        """Upload the momentum-exchange node and group tables (see
        ``force_tables``); returns a handle for ``momentum_exchange``."""
        nodes, grecs = force_tables(self.nx, self.ny, groups)
        nblocks = max(1, nodes.size // nt.FORCE_BLOCK)
        with torch.cuda.stream(self.stream):
            h = {
                "nodes": torch.from_numpy(nodes.view(np.uint8).copy()).to(
                    self.device),
                "groups": torch.from_numpy(grecs.view(np.uint8).copy()).to(
                    self.device),
                "n_nodes": int(nodes.size), "n_groups": int(grecs.size),
                "work": torch.zeros(3 * nblocks, dtype=torch.float64,
                                    device=self.device),
                "out": torch.zeros(3 * max(1, grecs.size),
                                   dtype=torch.float64, device=self.device),
            }
        return h

    def momentum_exchange(self, handle):
        """(n_groups, 3) FP64 array of force x, force y and torque on each
        group's links for the current step: f* is the current post-
        collision state, f_new its streamed and boundary-treated values
        (read when in memory, else recomputed for the nodes alone)."""
        fpop = self.f[self.i_pop]
        fnew = (nt.ptr(self.f[self.i_pop_new]) if self.i_pop_new is not None
                else None)
        nt.check(self.lib.d2q9_momentum_exchange(
            self._Lp, nt.ptr(fpop), fnew, nt.ptr(self.mom[self.i_mom]),
            nt.ptr(handle["nodes"]), handle["n_nodes"],
            nt.ptr(handle["groups"]), handle["n_groups"],
            nt.ptr(handle["work"]), nt.ptr(handle["out"]), self._stream_ptr),
            "d2q9_momentum_exchange")
        self.synchronize()
        return handle["out"][:3 * handle["n_groups"]].cpu().numpy().reshape(
            -1, 3)

    def residue_sums(self, mom_old):
        """Six FP64 residue sums of the current moments against
        ``mom_old`` (updated in place)."""
        if not self.mom_full:
            raise RuntimeError("moments of this step were not stored")
        nt.check(self.lib.d2q9_residues(
            self._Lp, nt.ptr(self.mom[self.i_mom]), nt.ptr(mom_old),
            nt.ptr(self.work), nt.ptr(self.red_out), self._stream_ptr),
            "d2q9_residues")
        self.synchronize()
        return self.red_out[:6].cpu().numpy().copy()

    # ---- host views (reference layouts) -----------------------------------
    def _decode_aos(self, idx):
        """Populations of buffer ``idx`` as a float64 numpy [size, 9]."""
        with torch.cuda.stream(self.stream):
            aos = pop_aos(self.f[idx], self.nx, self.ny, self.B,
                          self.nbj).to(torch.float64)
            if self.storage != nt.STORE_F64:
                w = torch.tensor(self.weights, dtype=torch.float64,
                                 device=aos.device)
                sc = torch.tensor([self.L.dev_scale[k] for k in range(9)],
                                  dtype=torch.float64, device=aos.device)
                aos = w + aos / sc
            aos = aos.cpu().numpy()
        if self.storage != nt.STORE_F64:
            aos[~self.tables.fluid_pad.ravel()] = 0.0
        return aos

    def pop(self):
        """Reference ``pop_fluid`` (post-collision) as float64 [size, 9]."""
        return self._decode_aos(self.i_pop)

    def pop_new(self):
        """Reference ``pop_fluid_new`` as float64 [size, 9]."""
        return self._decode_aos(self.materialize())

    def moments_host(self):
        """(density [size], velocity [size, 2]) of the current step."""
        if not self.mom_full:
            raise RuntimeError(
                "moments of this step were not stored: advance with "
                "write_moments=True")
        with torch.cuda.stream(self.stream):
            m = self.mom[self.i_mom]
            rho = self._grid(m, 0).to(torch.float64).reshape(-1).cpu().numpy()
            ux = self._grid(m, 1).to(torch.float64).reshape(-1).cpu().numpy()
            uy = self._grid(m, 2).to(torch.float64).reshape(-1).cpu().numpy()
        return rho, np.stack([ux, uy], axis=-1)
