// Tile and cell sweeps shared by the single-lattice kernels
// (d2q9_step.hip) and the batched persistent kernel (d2q9_batched.hip).
// See d2q9_step.hip for the launch structure.
#pragma once

#include "d2q9_device.h"

namespace d2q9 {

enum { PHASE_COLLIDE = 0, PHASE_FUSED = 1, PHASE_STREAM = 2 };

// One lane's share of one direction: CPL raw elements moved as a single
// 4- or 8-byte word.
template <int BYTES> struct LaneWord;
template <> struct LaneWord<4> { using type = uint32_t; };
template <> struct LaneWord<8> { using type = uint64_t; };

template <typename R, int CPL>
struct alignas(sizeof(R) * CPL) Vec8 {
  R v[CPL];
};

template <typename R, int CPL>
__device__ __forceinline__ void store8_nt(R* p, const Vec8<R, CPL>& v) {
  using W = typename LaneWord<sizeof(R) * CPL>::type;
  W w;
  __builtin_memcpy(&w, &v, sizeof(W));
  __builtin_nontemporal_store(w, reinterpret_cast<W*>(p));
}

// Pre-collision values of one flagged cell: bounce-back off solid
// neighbours, periodic wrap and boundary rules (streaming_kernels.py:29-47,
// MPI_kernels.py:34-99, fluid_boundary_kernels.py:5-152).
template <int S, int PHASE>
__device__ __forceinline__ void gather_cell(
    const d2q9_lattice& L, const Consts<typename Codec<S>::T>& K,
    const typename Codec<S>::raw* __restrict__ fsrc,
    const typename Codec<S>::T* __restrict__ msrc, int i, int j, int64_t c,
    uint8_t fl, typename Codec<S>::T (&g)[9]) {
  if (PHASE == PHASE_COLLIDE) {
#pragma unroll
    for (int k = 0; k < 9; ++k) g[k] = ld<S>(L, K, fsrc, i, j, k);
  } else {
    g[0] = ld<S>(L, K, fsrc, i, j, 0);
#pragma unroll
    for (int k = 1; k < 9; ++k) g[k] = pull<S>(L, K, fsrc, msrc, i, j, c, fl, k);
    if (fl & D2Q9_CELL_BC) apply_bc<S>(L, K, fsrc, msrc, i, j, c, g);
  }
}

// Coalesced gather of one lane's CPL cells (j0 .. j0+CPL-1) of row i: the
// nine directions of block tj+1 of the pull rows.  The y-shift is a shifted
// address when CPL == 1; for CPL > 1 it is a cross-lane shuffle, and the two
// cells at the ends of the wave's block take their neighbour from one
// six-lane "halo" load (lanes 0..5 fetch, for directions 2,5,6 the element
// before the block and for 4,7,8 the element after it) broadcast with
// readlane.  All ten loads are issued before any is consumed.  Exact for
// interior cells away from periodic edges; no address leaves the padded row.
__device__ __forceinline__ uint32_t bits32(float v) { return __float_as_uint(v); }
__device__ __forceinline__ float from_bits32(uint32_t v) { return __uint_as_float(v); }

template <typename T>
__device__ __forceinline__ T wave_readlane(T v, int lane);
template <>
__device__ __forceinline__ float wave_readlane<float>(float v, int lane) {
  return __uint_as_float((uint32_t)__builtin_amdgcn_readlane((int)__float_as_uint(v), lane));
}

// Whole-wave shifts by one lane (GFX9 DPP wave_shr:1 / wave_shl:1): lane l
// receives the value of lane l - 1 (from_prev) or l + 1 (from_next); the
// lane without a source (0, resp. 63) keeps its own value, as __shfl_up /
// __shfl_down do.  One VALU move per dword instead of an LDS-crossbar
// permute and its wait.
template <int CTRL, typename T>
__device__ __forceinline__ T wave_shift(T v) {
  static_assert(sizeof(T) % 4 == 0, "dword-sized values");
  constexpr int N = sizeof(T) / 4;
  int w[N];
  __builtin_memcpy(w, &v, sizeof(T));
#pragma unroll
  for (int n = 0; n < N; ++n) w[n] = __builtin_amdgcn_update_dpp(w[n], w[n], CTRL, 0xf, 0xf, false);
  T r;
  __builtin_memcpy(&r, w, sizeof(T));
  return r;
}
template <typename T>
__device__ __forceinline__ T wave_from_prev(T v) { return wave_shift<0x138>(v); }
template <typename T>
__device__ __forceinline__ T wave_from_next(T v) { return wave_shift<0x130>(v); }

// Raw words of one lane's gather (one per direction) and, for CPL > 1,
// the halo element; loaded by gather_issue, decoded by gather_finish.
template <int S>
struct RawBlock {
  using W = typename LaneWord<sizeof(typename Codec<S>::raw) * Codec<S>::CPL>::type;
  W w[9];
  typename Codec<S>::raw halo;
};

// Row r of a lattice periodic in x, for r in [1 - nx, 2 nx]: the image in
// 1..nx (the two-step sweep's grown rows reach past the first and last row).
__device__ __forceinline__ int wrap_row(const d2q9_lattice& L, int r) {
  if (L.x_periodic) {
    if (r < 1) r += L.nx;
    else if (r > L.nx) r -= L.nx;
  }
  return r;
}

// WRAP: rows outside 1..nx are taken from their periodic image (the
// two-step sweep on a lattice periodic in x); otherwise row i - c_x is read
// as it is (0 and nx + 1 being ghost rows).
template <int S, int PHASE, bool WRAP = false>
__device__ __forceinline__ void gather_issue(const d2q9_lattice& L,
                                             const typename Codec<S>::raw* __restrict__ fsrc,
                                             int i, int tj, int lane, int j0, RawBlock<S>& rb) {
  using C = Codec<S>;
  using R = typename C::raw;
  constexpr int CPL = C::CPL;
  constexpr int B = 64 * CPL;
  auto rw = [&](int r) { return WRAP ? wrap_row(L, r) : r; };
  if constexpr (CPL == 1) {
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      const int row = rw((PHASE == PHASE_COLLIDE) ? i : i - CX[k]);
      const int cy = (PHASE == PHASE_COLLIDE) ? 0 : CY[k];
      const int jc = (j0 <= L.ny) ? j0 - cy : j0;
      rb.w[k] = *reinterpret_cast<const typename RawBlock<S>::W*>(
          fsrc + pop_index<B>(L, row, jc, k));
    }
  } else {
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      const int row = rw((PHASE == PHASE_COLLIDE) ? i : i - CX[k]);
      rb.w[k] = *reinterpret_cast<const typename RawBlock<S>::W*>(
          fsrc + (((int64_t)row * L.nbj + tj + 1) * 9 + k) * B + lane * CPL);
    }
    rb.halo = R();
    if (PHASE != PHASE_COLLIDE) {
      // lane h < 6 fetches the end element of direction HK[h]
      const int h = lane < 6 ? lane : 0;
      const int kh = h == 0 ? 2 : h == 1 ? 5 : h == 2 ? 6 : h == 3 ? 4 : h == 4 ? 7 : 8;
      const int cxh = h == 0 ? 0 : h == 1 ? 1 : h == 2 ? -1 : h == 3 ? 0 : h == 4 ? -1 : 1;
      const int jf = 1 + tj * B;             // first cell of the block
      int jh = h < 3 ? jf - 1 : jf + B;      // element before / after
      if (jh > L.ny + 1) jh = jf;            // past the padded row: unused
      rb.halo = fsrc[pop_index<B>(L, rw(i - cxh), jh, kh)];
    }
  }
}

template <int S, int PHASE>
__device__ __forceinline__ void gather_finish(const Consts<typename Codec<S>::T>& K,
                                              const RawBlock<S>& rb, int lane,
                                              typename Codec<S>::T (&g)[Codec<S>::CPL][9]) {
  using C = Codec<S>;
  using T = typename C::T;
  using R = typename C::raw;
  constexpr int CPL = C::CPL;
  if constexpr (CPL == 1) {
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      R r;
      __builtin_memcpy(&r, &rb.w[k], sizeof(R));
      g[0][k] = C::dec(r, k, K);
    }
  } else {
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      const int cy = (PHASE == PHASE_COLLIDE) ? 0 : CY[k];
      Vec8<R, CPL> v;
      __builtin_memcpy(&v, &rb.w[k], sizeof(v));
      T x[CPL];
#pragma unroll
      for (int q = 0; q < CPL; ++q) x[q] = C::dec(v.v[q], k, K);
      if (cy == 0) {
#pragma unroll
        for (int q = 0; q < CPL; ++q) g[q][k] = x[q];
      } else {
        const int h = k == 2 ? 0 : k == 5 ? 1 : k == 6 ? 2 : k == 4 ? 3 : k == 7 ? 4 : 5;
        const T hv = wave_readlane<T>(C::dec(rb.halo, k, K), h);
        if (cy == 1) {  // cell j takes j - 1
          T prev = wave_from_prev(x[CPL - 1]);
          g[0][k] = lane == 0 ? hv : prev;
#pragma unroll
          for (int q = 1; q < CPL; ++q) g[q][k] = x[q - 1];
        } else {  // cell j takes j + 1
          T next = wave_from_next(x[0]);
#pragma unroll
          for (int q = 0; q < CPL - 1; ++q) g[q][k] = x[q + 1];
          g[CPL - 1][k] = lane == 63 ? hv : next;
        }
      }
    }
  }
}

template <int S, int PHASE>
__device__ __forceinline__ void gather_block(
    const d2q9_lattice& L, const Consts<typename Codec<S>::T>& K,
    const typename Codec<S>::raw* __restrict__ fsrc, int i, int tj, int lane,
    int j0, typename Codec<S>::T (&g)[Codec<S>::CPL][9]) {
  RawBlock<S> rb;
  gather_issue<S, PHASE>(L, fsrc, i, tj, lane, j0, rb);
  gather_finish<S, PHASE>(K, rb, lane, g);
}

// Moments + collision (or nothing for PHASE_STREAM) of one cell, encoded
// into out[.][q]; rho/u stored when asked.  With ENC false the collided
// values stay in g and the caller encodes them (pairs of 16-bit cells).
template <int S, int MODEL, int FORCING, int PHASE, bool ENC = true>
__device__ __forceinline__ void finish_cell(
    const d2q9_lattice& L, const Consts<typename Codec<S>::T>& K,
    typename Codec<S>::T (&g)[9], typename Codec<S>::raw (&out)[9][Codec<S>::CPL],
    int q, int64_t c, bool store_mom, typename Codec<S>::T* __restrict__ mdst) {
  using C = Codec<S>;
  using T = typename C::T;
  if constexpr (PHASE == PHASE_STREAM) {
#pragma unroll
    for (int k = 0; k < 9; ++k) out[k][q] = C::enc(g[k], k, K);
  } else {
    T ffx = (T)0, ffy = (T)0;
    if (FORCING != D2Q9_FORCING_NONE && L.force_source == D2Q9_FSRC_FIELD) {
      const T* ff = static_cast<const T*>(L.force_field);
      ffx = ff[c];
      ffy = ff[L.plane + c];
    }
    T drho, rho, ux, uy, fx, fy, px, py;
    moments<T, FORCING, C::dev>(g, L, K, ffx, ffy, drho, rho, ux, uy, fx, fy, px, py);
    collide<T, MODEL, FORCING, C::dev>(g, K, drho, rho, ux, uy, fx, fy, px, py);
    if constexpr (ENC) {
      // 16-bit storage rounds stochastically (deterministic in the data)
      const uint32_t key = C::dev && sizeof(typename C::raw) == 2 ? cell_key((uint32_t)c) : 0u;
#pragma unroll
      for (int k = 0; k < 9; ++k) out[k][q] = C::encs(g[k], k, key, K);
    }
    if (store_mom) {
      mdst[c] = rho;
      mdst[L.plane + c] = ux;
      mdst[2 * L.plane + c] = uy;
    }
  }
}

template <int S>
__device__ __forceinline__ void store_block(
    const d2q9_lattice& L, typename Codec<S>::raw* __restrict__ fdst, int i,
    int tj, int lane, const typename Codec<S>::raw (&out)[9][Codec<S>::CPL]) {
  using R = typename Codec<S>::raw;
  constexpr int CPL = Codec<S>::CPL;
  constexpr int B = 64 * CPL;
  const int64_t obase = (((int64_t)i * L.nbj + tj + 1) * 9) * B + lane * CPL;
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    if constexpr (CPL == 1) {
      store_nt(out[k][0], fdst + obase + (int64_t)k * B);
    } else {
      Vec8<R, CPL> v;
#pragma unroll
      for (int q = 0; q < CPL; ++q) v.v[q] = out[k][q];
      store8_nt<R, CPL>(fdst + obase + (int64_t)k * B, v);
    }
  }
}

// Straight-wall cells inside a bulk tile (masks per row and side, uniform
// over the wave).  Zero-gradient (sides 0, 1): a marked cell takes the
// unknown directions (2,5,6 from the cell at j + 1, or 4,7,8 from the cell
// at j - 1) as that cell pulled them, before any boundary rule -- the same
// values the rule reads (apply_bc, zero-gradient branch) when the source
// cell has no solid neighbour and no periodic wrap, which the host checks;
// the source is the next cell of the lane or, at a lane edge, of the
// neighbouring lane (one shuffle).  Bounce-back (sides 2..5): the unknown
// directions take the cell's own post-collision value of the opposite
// direction (fluid_boundary_kernels.py:5-25), one load each, lines the
// tile's neighbours have just read.
template <int S, int PHASE>
__device__ __forceinline__ void wall_rows(const d2q9_lattice& L,
                                          const Consts<typename Codec<S>::T>& K,
                                          const typename Codec<S>::raw* __restrict__ fsrc,
                                          typename Codec<S>::T (&g)[Codec<S>::CPL][9], int slot,
                                          int lane, int i, int j0) {
  using T = typename Codec<S>::T;
  constexpr int CPL = Codec<S>::CPL;
  const int row = __builtin_amdgcn_readfirstlane((int)threadIdx.y);
  const uint64_t* zrow = L.wall_masks + ((int64_t)slot * blockDim.y + row) * 6 * CPL;
  bool any[6];
#pragma unroll
  for (int s = 0; s < 6; ++s) {
    any[s] = false;
#pragma unroll
    for (int w = 0; w < CPL; ++w) any[s] |= zrow[s * CPL + w] != 0;
  }
  auto lane_bits = [&](int s) {
    uint32_t b = 0;
#pragma unroll
    for (int q = 0; q < CPL; ++q) {
      const int o = lane * CPL + q;
      b |= (uint32_t)((zrow[s * CPL + (o >> 6)] >> (o & 63)) & 1ull) << q;
    }
    return b;
  };
  constexpr int DU[3] = {2, 5, 6};
  constexpr int DD[3] = {4, 7, 8};
  if (any[0]) {
    const uint32_t up = lane_bits(0);
#pragma unroll
    for (int n = 0; n < 3; ++n) {
      const int d = DU[n];
      const T nxt = wave_from_next(g[0][d]);
#pragma unroll
      for (int q = 0; q < CPL; ++q) {
        const T src = (q + 1 < CPL) ? g[q + 1][d] : nxt;
        if ((up >> q) & 1) g[q][d] = src;
      }
    }
  }
  if (any[1]) {
    const uint32_t down = lane_bits(1);
#pragma unroll
    for (int n = 0; n < 3; ++n) {
      const int d = DD[n];
      const T prv = wave_from_prev(g[CPL - 1][d]);
#pragma unroll
      for (int q = CPL - 1; q >= 0; --q) {
        const T src = (q > 0) ? g[q - 1][d] : prv;
        if ((down >> q) & 1) g[q][d] = src;
      }
    }
  }
  constexpr int BD[4][3] = {{2, 5, 6}, {4, 7, 8}, {1, 5, 8}, {3, 6, 7}};
#pragma unroll
  for (int s = 2; s < 6; ++s) {
    if (!any[s]) continue;
    const uint32_t bits = lane_bits(s);
#pragma unroll
    for (int q = 0; q < CPL; ++q) {
      if (!((bits >> q) & 1)) continue;
#pragma unroll
      for (int n = 0; n < 3; ++n) {
        const int d = BD[s - 2][n];
        g[q][d] = ld<S>(L, K, fsrc, i, j0 + q, INV[d]);
      }
    }
  }
}

// Bulk tile at position p of the tile list (host-classified: no solid
// cell).  Every cell is swept as plain interior fluid and no flag is read;
// the cells marked in bulk_masks are not stored here: the special ones
// (solid neighbour, boundary rule, stored moment, periodic edge) are
// written by the fix-up pass, the ones past the lattice's last column are
// never written.
// Position of bulk tile p for this wave: row i, block tj, first cell j0 of
// the lane, and the lane's masked cells.
struct BulkPos {
  int i, tj, j0;
  uint32_t skip;  // masked cells of the lane (bit q: cell j0 + q)
  int wslot;      // the tile's slot in the wall masks, or -1
};

// Row, block and first cell of the lane for bulk tile p.  Each wave sweeps
// one row of the tile (workgroup 64 x rows).
template <int S>
__device__ __forceinline__ BulkPos bulk_place(const d2q9_lattice& L, int ntj, int p) {
  constexpr int CPL = Codec<S>::CPL;
  constexpr int B = 64 * CPL;
  BulkPos bp;
  const int t = L.tiles[p];
  bp.tj = t % ntj;
  const int ti = t / ntj;
  bp.i = 1 + ti * (int)blockDim.y + (int)threadIdx.y;
  bp.j0 = 1 + bp.tj * B + (int)threadIdx.x * CPL;
  bp.skip = 0;
  bp.wslot = -1;
  return bp;
}

// The wave's mask words (uniform over the wave: scalar loads), issued
// before the gather so that their latency overlaps it.
template <int S>
__device__ __forceinline__ void bulk_mask_load(const d2q9_lattice& L, int p,
                                               uint64_t (&m)[Codec<S>::CPL]) {
  constexpr int CPL = Codec<S>::CPL;
  const int row = __builtin_amdgcn_readfirstlane((int)threadIdx.y);
  const uint64_t* mrow = L.bulk_masks + ((int64_t)p * blockDim.y + row) * CPL;
#pragma unroll
  for (int w = 0; w < CPL; ++w) m[w] = mrow[w];
}

template <int S>
__device__ __forceinline__ uint32_t bulk_skip(const uint64_t (&m)[Codec<S>::CPL], int lane) {
  constexpr int CPL = Codec<S>::CPL;
  uint32_t skip = 0;
#pragma unroll
  for (int q = 0; q < CPL; ++q) {
    const int o = lane * CPL + q;
    skip |= (uint32_t)((m[o >> 6] >> (o & 63)) & 1ull) << q;
  }
  return skip;
}

// Everything after the loads of one bulk tile: decode and shuffle, the
// zero-gradient rows, moments + collision, the stores.
template <int S, int MODEL, int FORCING, int PHASE>
__device__ __forceinline__ void bulk_finish(
    const d2q9_lattice& L, const Consts<typename Codec<S>::T>& K, const RawBlock<S>& rb,
    const BulkPos& bp, const typename Codec<S>::raw* __restrict__ fsrc,
    typename Codec<S>::raw* __restrict__ fdst,
    typename Codec<S>::T* __restrict__ mdst, int write_mom, int p) {
  using C = Codec<S>;
  using T = typename C::T;
  constexpr int CPL = C::CPL;
  constexpr int B = 64 * CPL;
  const int lane = (int)threadIdx.x;
  T g[CPL][9];
  gather_finish<S, PHASE>(K, rb, lane, g);
  if constexpr (PHASE != PHASE_COLLIDE) {
    if (bp.wslot >= 0) wall_rows<S, PHASE>(L, K, fsrc, g, bp.wslot, lane, bp.i, bp.j0);
  }
  typename C::raw out[9][CPL];
  constexpr bool PAIR = C::dev && sizeof(typename C::raw) == 2 && CPL == 2 &&
                        PHASE != PHASE_STREAM;
  if constexpr (PAIR) {
    // two cells per lane: collide both, then round and pack them together
    const int64_t c0 = cell_index(L, bp.i, bp.j0);
#pragma unroll
    for (int q = 0; q < CPL; ++q)
      finish_cell<S, MODEL, FORCING, PHASE, false>(L, K, g[q], out, q, c0 + q,
                                                   write_mom != 0 && !((bp.skip >> q) & 1),
                                                   mdst);
    const uint32_t key0 = cell_key((uint32_t)c0), key1 = cell_key((uint32_t)(c0 + 1));
    uint32_t w[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) w[k] = C::encs2(g[0][k], g[1][k], k, key0, key1, K);
    if (bp.skip == 0) {
      const int64_t obase = (((int64_t)bp.i * L.nbj + bp.tj + 1) * 9) * B + lane * CPL;
#pragma unroll
      for (int k = 0; k < 9; ++k)
        __builtin_nontemporal_store(w[k],
                                    reinterpret_cast<uint32_t*>(fdst + obase + (int64_t)k * B));
      return;
    }
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      const unsigned short h0 = (unsigned short)(w[k] & 0xffffu);
      const unsigned short h1 = (unsigned short)(w[k] >> 16);
      __builtin_memcpy(&out[k][0], &h0, 2);
      __builtin_memcpy(&out[k][1], &h1, 2);
    }
  } else {
#pragma unroll
    for (int q = 0; q < CPL; ++q)
      finish_cell<S, MODEL, FORCING, PHASE>(L, K, g[q], out, q, cell_index(L, bp.i, bp.j0 + q),
                                            write_mom != 0 && !((bp.skip >> q) & 1), mdst);
  }
  if (bp.skip == 0) {
    store_block<S>(L, fdst, bp.i, bp.tj, lane, out);
  } else {
#pragma unroll
    for (int q = 0; q < CPL; ++q) {
      if ((bp.skip >> q) & 1) continue;
#pragma unroll
      for (int k = 0; k < 9; ++k) fdst[pop_index<B>(L, bp.i, bp.j0 + q, k)] = out[k][q];
    }
  }
}

// Bulk tile at position p of the tile list (host-classified: no solid
// cell).  Every cell is swept as plain interior fluid and no flag is read;
// the cells marked in bulk_masks are not stored here: the special ones
// (solid neighbour, boundary rule, stored moment, periodic edge) are
// written by the fix-up pass, the ones past the lattice's last column are
// never written.
template <int S, int MODEL, int FORCING, int PHASE>
__device__ __forceinline__ void bulk_tile(
    const d2q9_lattice& L, const typename Codec<S>::raw* __restrict__ fsrc,
    typename Codec<S>::raw* __restrict__ fdst,
    typename Codec<S>::T* __restrict__ mdst, int write_mom, int ntj, int p) {
  using T = typename Codec<S>::T;
  BulkPos bp = bulk_place<S>(L, ntj, p);
  if (bp.i > L.nx) return;  // whole wave (last tile row of a partial tile)
  // tile tables first, then the gather: none of the loads waits on another
  uint64_t m[Codec<S>::CPL];
  bulk_mask_load<S>(L, p, m);
  if (PHASE != PHASE_COLLIDE && L.wall_masks) bp.wslot = L.wall_index[p];
  const Consts<T> K = load_consts<T>(L);
  RawBlock<S> rb;
  gather_issue<S, PHASE>(L, fsrc, bp.i, bp.tj, (int)threadIdx.x, bp.j0, rb);
  bp.skip = bulk_skip<S>(m, (int)threadIdx.x);
  bulk_finish<S, MODEL, FORCING, PHASE>(L, K, rb, bp, fsrc, fdst, mdst, write_mom, p);
}

// Fix-up cells [first, first + workgroup size) of the lattice's list, one
// thread per cell.
template <int S, int MODEL, int FORCING, int PHASE>
__device__ __forceinline__ void fix_chunk(
    const d2q9_lattice& L, const typename Codec<S>::raw* __restrict__ fsrc,
    typename Codec<S>::raw* __restrict__ fdst,
    const typename Codec<S>::T* __restrict__ msrc,
    typename Codec<S>::T* __restrict__ mdst, int write_mom, int first) {
  using C = Codec<S>;
  using T = typename C::T;
  constexpr int CPL = C::CPL;
  constexpr int B = 64 * CPL;
  const int n = first + (int)(threadIdx.y * blockDim.x + threadIdx.x);
  if (n >= L.n_fix) return;
  const Consts<T> K = load_consts<T>(L);
  const int64_t c = L.fix_cells[n];
  const int i = (int)((c - L.origin) / L.pitch), j = (int)((c - L.origin) % L.pitch);
  const uint8_t fl = L.flags[c];
  T g[9];
  gather_cell<S, PHASE>(L, K, fsrc, msrc, i, j, c, fl, g);
  typename C::raw out[9][CPL];
  finish_cell<S, MODEL, FORCING, PHASE>(
      L, K, g, out, 0, c, write_mom || (fl & D2Q9_CELL_STORE_MOM), mdst);
#pragma unroll
  for (int k = 0; k < 9; ++k) fdst[pop_index<B>(L, i, j, k)] = out[k][0];
}

// General tile t (one holding solid cells or reaching past the lattice):
// coalesced gather for every lane, the special cells re-gathered one by
// one, non-fluid cells written back as the encoded zero.
template <int S, int MODEL, int FORCING, int PHASE>
__device__ __forceinline__ void general_tile(
    const d2q9_lattice& L, const typename Codec<S>::raw* __restrict__ fsrc,
    typename Codec<S>::raw* __restrict__ fdst,
    const typename Codec<S>::T* __restrict__ msrc,
    typename Codec<S>::T* __restrict__ mdst, int write_mom, int ntj, int t) {
  using C = Codec<S>;
  using T = typename C::T;
  constexpr int CPL = C::CPL;
  constexpr int B = 64 * CPL;
  const int tj = t % ntj, ti = t / ntj;
  const int lane = (int)threadIdx.x;
  const int i = 1 + ti * (int)blockDim.y + (int)threadIdx.y;
  const int j0 = 1 + tj * B + lane * CPL;
  if (i > L.nx) return;  // whole wave
  const Consts<T> K = load_consts<T>(L);
  T g[CPL][9];
  gather_block<S, PHASE>(L, K, fsrc, i, tj, lane, j0, g);
  uint8_t fl[CPL];
#pragma unroll
  for (int q = 0; q < CPL; ++q) {
    const int j = j0 + q;
    fl[q] = 0;
    if (j > L.ny) continue;
    const int64_t c = cell_index(L, i, j);
    fl[q] = L.flags[c];
    if (!(fl[q] & D2Q9_CELL_FLUID)) continue;
    const bool edge = (L.x_periodic && (i == 1 || i == L.nx)) ||
                      (L.y_periodic && (j == 1 || j == L.ny));
    if (fl[q] != D2Q9_CELL_FLUID || (edge && PHASE != PHASE_COLLIDE))
      gather_cell<S, PHASE>(L, K, fsrc, msrc, i, j, c, fl[q], g[q]);
  }
  typename C::raw out[9][CPL];
#pragma unroll
  for (int q = 0; q < CPL; ++q) {
    if (!(fl[q] & D2Q9_CELL_FLUID)) {
#pragma unroll
      for (int k = 0; k < 9; ++k) out[k][q] = enc_zero<S>(k, K);
      continue;
    }
    finish_cell<S, MODEL, FORCING, PHASE>(
        L, K, g[q], out, q, cell_index(L, i, j0 + q),
        write_mom || (fl[q] & D2Q9_CELL_STORE_MOM), mdst);
  }
  // Cells beyond ny in a partial block: the ghost column ny+1 gets its
  // encoded zero back, the block padding after it is never read.
  if (CPL == 1 && j0 > L.ny + 1) return;
  store_block<S>(L, fdst, i, tj, lane, out);
}

}  // namespace d2q9
