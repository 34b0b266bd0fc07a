// Single-lattice D2Q9 sweep kernels and their C ABI (include/d2q9.h).
//
// A wave owns one AoSoA block (B = 64 * CPL cells along y) of one row; a
// workgroup stacks tile_i such rows.  Each lane moves one word per
// direction (1 double, 2 floats or 2 halves), reads the nine directions at
// their pull offsets (the y-shift crosses lanes by a shuffle when CPL > 1)
// and writes them back unshifted with non-temporal stores, so every HBM
// access of a wave is one contiguous segment and the nine directions of a
// block sit side by side.
//
// Bulk tiles (no solid cell) sweep every cell as plain interior fluid and
// never read the flag array; straight-wall rows are done in the tile from
// per-tile masks, and the other special cells (solid neighbour, boundary
// rule, periodic edge, a moment the next step reads) are masked out of the
// store and recomputed one per thread by the fix-up pass, with their flag
// byte.  General tiles (holding a solid) handle each cell with its flag
// byte.  FP64 runs the bulk tiles and the general/fix-up pass as two
// overlapped kernels; the float forms run a fused step as one kernel.
#include <hip/hip_runtime.h>

#include <cstddef>
#include <initializer_list>
#include <map>
#include <memory>
#include <mutex>
#include <utility>
#include <cstdio>
#include <cstring>
#include <string>

#include "d2q9_device.h"
#include "d2q9_reduce.h"
#include "d2q9_sweep.h"

using namespace d2q9;

namespace {
thread_local std::string g_last_error;
}  // namespace

void d2q9::set_error(const char* what) { g_last_error = what; }

namespace {

int fail(int code, const char* what) {
  g_last_error = what;
  return code;
}

int check_hip(hipError_t e, const char* where) {
  if (e == hipSuccess) return 0;
  g_last_error = std::string(where) + ": " + hipGetErrorString(e);
  return (int)e;
}

template <int S, int MODEL, int FORCING, int PHASE>
__global__ __launch_bounds__(256) void k_bulk(
    const d2q9_lattice L, const typename Codec<S>::raw* __restrict__ fsrc,
    typename Codec<S>::raw* __restrict__ fdst,
    typename Codec<S>::T* __restrict__ mdst, int write_mom, int ntj) {
  bulk_tile<S, MODEL, FORCING, PHASE>(L, fsrc, fdst, mdst, write_mom, ntj,
                                      xcd_remap(blockIdx.x, L.n_bulk));
}

// Second pass.  Workgroups [0, ngen) sweep the general tiles, the ones after
// that recompute the fix-up cells of bulk tiles over what k_bulk stored
// (the two sets of cells are disjoint, so the passes may overlap).
template <int S, int MODEL, int FORCING, int PHASE>
__global__ __launch_bounds__(256) void k_general_fix(
    const d2q9_lattice L, const typename Codec<S>::raw* __restrict__ fsrc,
    typename Codec<S>::raw* __restrict__ fdst,
    const typename Codec<S>::T* __restrict__ msrc,
    typename Codec<S>::T* __restrict__ mdst, int write_mom, int ntj, int ngen,
    int first) {
  if ((int)blockIdx.x >= ngen) {
    fix_chunk<S, MODEL, FORCING, PHASE>(L, fsrc, fdst, msrc, mdst, write_mom,
                                        ((int)blockIdx.x - ngen) * 64 * (int)blockDim.y);
    return;
  }
  const int r = xcd_remap(blockIdx.x, ngen);
  const int t = L.tiles ? L.tiles[first + r] : r;
  general_tile<S, MODEL, FORCING, PHASE>(L, fsrc, fdst, msrc, mdst, write_mom, ntj, t);
}

// One kernel for the whole step: workgroups [0, ngen) sweep the general
// tiles, the next nfixb the fix-up chunks, the rest the bulk tiles.  With
// GEN false the lattice has no general tile (their fluid cells are in the
// fix-up list) and the general path is not compiled in, which keeps the
// kernel at the bulk sweep's register count.
template <int S, int MODEL, int FORCING, int PHASE, bool GEN>
__global__ __launch_bounds__(256) void k_step_all(
    const d2q9_lattice L, const typename Codec<S>::raw* __restrict__ fsrc,
    typename Codec<S>::raw* __restrict__ fdst,
    const typename Codec<S>::T* __restrict__ msrc,
    typename Codec<S>::T* __restrict__ mdst, int write_mom, int ntj, int nfixb) {
  const int b = (int)blockIdx.x;
  const int ngen = GEN ? L.n_general : 0;
  if (GEN && b < ngen) {
    general_tile<S, MODEL, FORCING, PHASE>(L, fsrc, fdst, msrc, mdst, write_mom, ntj,
                                           L.tiles[L.n_bulk + b]);
  } else if (b < ngen + nfixb) {
    fix_chunk<S, MODEL, FORCING, PHASE>(L, fsrc, fdst, msrc, mdst, write_mom,
                                        (b - ngen) * 64 * (int)blockDim.y);
  } else {
    bulk_tile<S, MODEL, FORCING, PHASE>(L, fsrc, fdst, mdst, write_mom, ntj,
                                        xcd_remap(b - ngen - nfixb, L.n_bulk));
  }
}

// Two fused steps of a deep tile: DEEP_ROWS rows x one block (B = 64 * CPL
// cells) whose neighbourhood two cells deep holds only plain fluid cells
// (the host checks it), but for a straight zero-gradient or bounce-back
// wall in the lattice's first or last column (deep_wall_global and the
// wall case of stage two apply its rule; the ghost column beside it is not
// computed, since only the wall's replaced directions would read it).
// Stage one computes step n + 1 (pull from the
// lattice, moments, collision) on the tile grown by one cell on every side
// into LDS; stage two pulls step n + 1 from LDS and stores step n + 2.  Per
// cell and two steps the lattice is read about 1.2 times (the grown rows
// and columns, mostly L2 hits) and written once, against twice each for
// two single steps.  Waves 0 .. DEEP_ROWS + 1 sweep the grown rows with the
// single step's gather (lane l: cells j0 + CPL l ..); the last two waves
// compute the grown columns, j0 - 1 and j0 + B, of every grown row (lane
// r: row r).  Two waves of 14 scattered cells each finish stage one sooner
// than one wave of 28 (measured: FP16 and BF16 steps 8 % faster, FP64
// 0.5 %), and 16 waves still let two workgroups fill a CU.  Arithmetic per
// cell is that of the single step, so in FP64
// and FP32 storage the result is bitwise that of two single steps; in the
// 16-bit forms the intermediate step stays in FP32 (LDS) instead of being
// rounded to storage, and only the stored step is rounded.
// Build options (development variants): the tile height and the waves per
// SIMD the deep kernel is held to.
#ifndef D2Q9_DEEP_ROWS
#define D2Q9_DEEP_ROWS 12
#endif
#ifndef D2Q9_DEEP_WAVES_PER_SIMD
#define D2Q9_DEEP_WAVES_PER_SIMD 8
#endif
constexpr int DEEP_ROWS = D2Q9_DEEP_ROWS;
constexpr int DEEP_R1 = DEEP_ROWS + 2;
constexpr int DEEP_WAVES = DEEP_R1 + 2;

template <int S, int MODEL, int FORCING>
__device__ __forceinline__ void deep_collide(const d2q9_lattice& L,
                                             const Consts<typename Codec<S>::T>& K,
                                             typename Codec<S>::T (&g)[9], int64_t c,
                                             typename Codec<S>::T& rho,
                                             typename Codec<S>::T& ux,
                                             typename Codec<S>::T& uy) {
  using T = typename Codec<S>::T;
  T ffx = (T)0, ffy = (T)0;
  if (FORCING != D2Q9_FORCING_NONE && L.force_source == D2Q9_FSRC_FIELD) {
    const T* ff = static_cast<const T*>(L.force_field);
    ffx = ff[c];
    ffy = ff[L.plane + c];
  }
  T drho, fx, fy, px, py;
  moments<T, FORCING, Codec<S>::dev>(g, L, K, ffx, ffy, drho, rho, ux, uy, fx, fy, px, py);
  collide<T, MODEL, FORCING, Codec<S>::dev>(g, K, drho, rho, ux, uy, fx, fy, px, py);
}

// Wall kinds of a deep tile at the low-y (j = 1) and high-y (j = ny) walls:
// 0 none, 1 zero-gradient, 2 bounce-back.  Each wall cell's unknown
// directions (2,5,6 at j = 1; 4,7,8 at j = ny) take, for zero-gradient, the
// value its inner neighbour pulled (which lies at the wall cell's own
// column) and, for bounce-back, the cell's own post-collision value of the
// opposite direction: the rules of the bulk sweep's wall rows.
template <int S, bool WRAP>
__device__ __forceinline__ void deep_wall_global(const d2q9_lattice& L,
                                                 const Consts<typename Codec<S>::T>& K,
                                                 const typename Codec<S>::raw* __restrict__ fsrc,
                                                 typename Codec<S>::T (&g)[9], int kind, int i,
                                                 int j, int cy) {
#pragma unroll
  for (int d = 1; d < 9; ++d) {
    if (CY[d] != cy) continue;
    g[d] = kind == 1 ? ld<S>(L, K, fsrc, WRAP ? wrap_row(L, i - CX[d]) : i - CX[d], j, d)
                     : ld<S>(L, K, fsrc, WRAP ? wrap_row(L, i) : i, j, INV[d]);
  }
}

// Two workgroups (32 waves) per CU need 8 waves per SIMD, i.e. at most 64
// VGPRs, and every form is held to that: BGK takes 42-61 VGPRs, the
// reduced MRT 38-54.
//
// Whether the wave index is taken as wave-uniform (readfirstlane), so that
// the rows' addresses are formed once per wave in scalar registers instead
// of per lane: that removes 13-20 % of the deep kernel's vector
// instructions, and measured (same box, identical results) FP64 BGK
// cylinder 8192x2048 -0.7 %, FP16 -1.5 %, the 4096x1024 MRT + Guo channel
// -0.5 % (FP64) and -3 % (FP16, BF16), but FP32 BGK +1.1 % and BF16 BGK
// +2.6 %; those two keep the per-lane form.
template <int S, int MODEL>
constexpr bool deep_scalar_rows() {
  return !(MODEL == D2Q9_BGK && (S == D2Q9_STORE_F32 || S == D2Q9_STORE_BF16DEV));
}

// Stage one of k_deep2 (see there): the grown rows (waves 0 .. DEEP_R1 - 1)
// and the two grown columns (the last two waves) of step n + 1, into LDS.
// WRAP: the tile reaches past the first or last row of a lattice periodic
// in x, whose rows are then read from their periodic images; only such
// tiles pay for the wrap.
template <int S, int MODEL, int FORCING, bool WRAP>
__device__ __forceinline__ void deep_stage1(
    const d2q9_lattice& L, const Consts<typename Codec<S>::T>& K,
    const typename Codec<S>::raw* __restrict__ fsrc,
    typename Codec<S>::T (&s1)[DEEP_R1][9][64 * Codec<S>::CPL + 2], int w, int lane, int i0,
    int j0, int tj, int lo, int hi) {
  using T = typename Codec<S>::T;
  constexpr int CPL = Codec<S>::CPL;
  constexpr int B = 64 * CPL;
  if (w < DEEP_R1) {
    // grown row i0 - 1 + w, the block's B cells; the loads issue at raised
    // wave priority, so that they reach memory ahead of other waves'
    // arithmetic (measured: 0.7 % on the FP64 step)
    const int i = i0 - 1 + w, jl = j0 + lane * CPL;
    T g[CPL][9];
    __builtin_amdgcn_s_setprio(2);
    RawBlock<S> rb;
    gather_issue<S, PHASE_FUSED, WRAP>(L, fsrc, i, tj, lane, jl, rb);
    __builtin_amdgcn_s_setprio(0);
    gather_finish<S, PHASE_FUSED>(K, rb, lane, g);
    if (lo && lane == 0) deep_wall_global<S, WRAP>(L, K, fsrc, g[0], lo, i, 1, 1);
    if (hi && lane == 63) deep_wall_global<S, WRAP>(L, K, fsrc, g[CPL - 1], hi, i, L.ny, -1);
    const int64_t c = cell_index(L, WRAP ? wrap_row(L, i) : i, jl);
#pragma unroll
    for (int q = 0; q < CPL; ++q) {
      T rho, ux, uy;
      deep_collide<S, MODEL, FORCING>(L, K, g[q], c + q, rho, ux, uy);
#pragma unroll
      for (int k = 0; k < 9; ++k) s1[w][k][1 + lane * CPL + q] = g[q][k];
    }
  } else if (lane < DEEP_R1 && !((w - DEEP_R1) ? hi : lo)) {
    // grown column j0 - 1 (wave DEEP_R1) or j0 + B (wave DEEP_R1 + 1), not
    // the ghost column beside a wall
    const int r = lane, cl = (w - DEEP_R1) ? B + 1 : 0;
    const int i = i0 - 1 + r, j = j0 - 1 + cl;
    T g[9];
#pragma unroll
    for (int k = 0; k < 9; ++k)
      g[k] = ld<S>(L, K, fsrc, WRAP ? wrap_row(L, i - CX[k]) : i - CX[k], j - CY[k], k);
    T rho, ux, uy;
    deep_collide<S, MODEL, FORCING>(L, K, g, cell_index(L, WRAP ? wrap_row(L, i) : i, j), rho,
                                    ux, uy);
#pragma unroll
    for (int k = 0; k < 9; ++k) s1[r][k][cl] = g[k];
  }
}

// XW: the lattice is periodic in x; the kernel then holds the stage-one
// variant that reads rows past the ends from their images, taken by the
// tiles next to them.  (Compiled into the kernel of every lattice, the
// second variant had slowed the FP16 cylinder 8192x2048 step by 3 %, though
// none of its tiles takes it: a kernel of its own for periodic lattices.)
template <int S, int MODEL, int FORCING, bool XW>
__global__ __launch_bounds__(64 * DEEP_WAVES, D2Q9_DEEP_WAVES_PER_SIMD) void k_deep2(
    const d2q9_lattice L, const int32_t* __restrict__ deep, const uint8_t* __restrict__ walls,
    int n_deep, int ntj, const typename Codec<S>::raw* __restrict__ fsrc,
    typename Codec<S>::raw* __restrict__ fdst, typename Codec<S>::T* __restrict__ mdst,
    int write_mom) {
  using C = Codec<S>;
  using T = typename C::T;
  using R = typename C::raw;
  constexpr int CPL = C::CPL;
  constexpr int B = 64 * CPL;
  __shared__ T s1[DEEP_R1][9][B + 2];
  const int lane = (int)threadIdx.x;
  const int w = deep_scalar_rows<S, MODEL>() ? __builtin_amdgcn_readfirstlane((int)threadIdx.y)
                                             : (int)threadIdx.y;
  const int p = xcd_remap((int)blockIdx.x, n_deep);
  const int t = deep[p];
  const int wf = walls ? (int)walls[p] : 0;
  const int lo = wf & 3, hi = (wf >> 2) & 3;
  const int tj = t % ntj;
  // in a lattice periodic in x, the last row of tiles of a lattice whose
  // rows are not a multiple of DEEP_ROWS ends at row nx, overlapping the
  // row of tiles before it (cells swept by two tiles are written twice with
  // the same bits)
  const int i0 = XW ? min(1 + (t / ntj) * DEEP_ROWS, L.nx - DEEP_ROWS + 1)
                    : 1 + (t / ntj) * DEEP_ROWS;
  const int j0 = 1 + tj * B;
  // the whole workgroup leaves (before its barrier) for a tile whose grown
  // region is not inside the lattice -- but for the ghost column beside a
  // wall, and for rows past the first and last of a lattice periodic in x
  // (read from their periodic images; at least 2 DEEP_ROWS rows): every
  // access then stays in bounds
  if (t < 0 ||
      (XW ? (!L.x_periodic || L.nx < 2 * DEEP_ROWS || i0 < 1)
          : (i0 - 1 < 1 || i0 + DEEP_ROWS > L.nx)) ||
      (j0 - 1 < 1 && !(lo && j0 == 1)) || (j0 + B > L.ny && !(hi && j0 + B - 1 == L.ny)))
    return;
  const Consts<T> K = load_consts<T>(L);
  // rows past the lattice's ends are read only by a tile of a lattice
  // periodic in x whose pulls reach row 0 or nx + 1
  if constexpr (XW) {
    if (i0 < 3 || i0 + DEEP_ROWS + 1 > L.nx)
      deep_stage1<S, MODEL, FORCING, true>(L, K, fsrc, s1, w, lane, i0, j0, tj, lo, hi);
    else
      deep_stage1<S, MODEL, FORCING, false>(L, K, fsrc, s1, w, lane, i0, j0, tj, lo, hi);
  } else {
  // The same statements as deep_stage1<..., false>, kept inline: the
  // compiler then schedules this kernel exactly as it did before the
  // periodic variant existed (through the function, the FP16 cylinder
  // 8192x2048 step measured 2.4 % slower).
  if (w < DEEP_R1) {
    // grown row i0 - 1 + w, the block's B cells; the loads issue at raised
    // wave priority, so that they reach memory ahead of other waves'
    // arithmetic (measured: 0.7 % on the FP64 step)
    const int i = i0 - 1 + w, jl = j0 + lane * CPL;
    T g[CPL][9];
    __builtin_amdgcn_s_setprio(2);
    RawBlock<S> rb;
    gather_issue<S, PHASE_FUSED>(L, fsrc, i, tj, lane, jl, rb);
    __builtin_amdgcn_s_setprio(0);
    gather_finish<S, PHASE_FUSED>(K, rb, lane, g);
    if (lo && lane == 0) deep_wall_global<S, false>(L, K, fsrc, g[0], lo, i, 1, 1);
    if (hi && lane == 63) deep_wall_global<S, false>(L, K, fsrc, g[CPL - 1], hi, i, L.ny, -1);
    const int64_t c = cell_index(L, i, jl);
#pragma unroll
    for (int q = 0; q < CPL; ++q) {
      T rho, ux, uy;
      deep_collide<S, MODEL, FORCING>(L, K, g[q], c + q, rho, ux, uy);
#pragma unroll
      for (int k = 0; k < 9; ++k) s1[w][k][1 + lane * CPL + q] = g[q][k];
    }
  } else if (lane < DEEP_R1 && !((w - DEEP_R1) ? hi : lo)) {
    // grown column j0 - 1 (wave DEEP_R1) or j0 + B (wave DEEP_R1 + 1), not
    // the ghost column beside a wall
    const int r = lane, cl = (w - DEEP_R1) ? B + 1 : 0;
    const int i = i0 - 1 + r, j = j0 - 1 + cl;
    T g[9];
#pragma unroll
    for (int k = 0; k < 9; ++k)
      g[k] = ld<S>(L, K, fsrc, i - CX[k], j - CY[k], k);
    T rho, ux, uy;
    deep_collide<S, MODEL, FORCING>(L, K, g, cell_index(L, i, j), rho,
                                    ux, uy);
#pragma unroll
    for (int k = 0; k < 9; ++k) s1[r][k][cl] = g[k];
  }
  }
  __syncthreads();
  if (w >= 1 && w <= DEEP_ROWS) {
    const int i = i0 - 1 + w, jl = j0 + lane * CPL;
    const int64_t c = cell_index(L, i, jl);
    T g[CPL][9];
#pragma unroll
    for (int q = 0; q < CPL; ++q) {
      const int col = 1 + lane * CPL + q;
#pragma unroll
      for (int k = 0; k < 9; ++k) g[q][k] = s1[w - CX[k]][k][col - CY[k]];
      const int wk = (lo && lane == 0 && q == 0) ? lo : 0;
      const int wkh = (hi && lane == 63 && q == CPL - 1) ? hi : 0;
      if (wk | wkh) {
        const int kind = wk ? wk : wkh, cyw = wk ? 1 : -1;
#pragma unroll
        for (int d = 1; d < 9; ++d) {
          if (CY[d] != cyw) continue;
          g[q][d] = kind == 1 ? s1[w - CX[d]][d][col] : s1[w][INV[d]][col];
        }
      }
      T rho, ux, uy;
      deep_collide<S, MODEL, FORCING>(L, K, g[q], c + q, rho, ux, uy);
      if (write_mom) {
        mdst[c + q] = rho;
        mdst[L.plane + c + q] = ux;
        mdst[2 * L.plane + c + q] = uy;
      }
    }
    if constexpr (C::dev && sizeof(R) == 2 && CPL == 2) {
      // the lane's two cells rounded and packed together, as in the bulk
      // sweep
      const uint32_t key0 = cell_key((uint32_t)c), key1 = cell_key((uint32_t)(c + 1));
      const int64_t o = (((int64_t)i * L.nbj + tj + 1) * 9) * B + lane * CPL;
#pragma unroll
      for (int k = 0; k < 9; ++k)
        __builtin_nontemporal_store(C::encs2(g[0][k], g[1][k], k, key0, key1, K),
                                    reinterpret_cast<uint32_t*>(fdst + o + (int64_t)k * B));
    } else {
      R out[9][CPL];
#pragma unroll
      for (int q = 0; q < CPL; ++q)
#pragma unroll
        for (int k = 0; k < 9; ++k) out[k][q] = C::enc(g[q][k], k, K);
      store_block<S>(L, fdst, i, tj, lane, out);
    }
  }
}

// f = feq(rho0, u0) on fluid cells, the encoded zero on every other padded
// cell, in both buffers.
template <int S>
__global__ __launch_bounds__(256) void k_init(const d2q9_lattice L,
                                              typename Codec<S>::raw* fa,
                                              typename Codec<S>::raw* fb,
                                              const typename Codec<S>::T* mom) {
  using C = Codec<S>;
  using T = typename C::T;
  constexpr int B = 64 * C::CPL;
  const int64_t n = (int64_t)(L.nx + 2) * (L.ny + 2);
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const int i = (int)(e / (L.ny + 2)), j = (int)(e % (L.ny + 2));
  const int64_t c = cell_index(L, i, j);
  const int64_t P = L.plane;
  const Consts<T> K = load_consts<T>(L);
  if (L.flags[c] & D2Q9_CELL_FLUID) {
    const T rho = mom[c], ux = mom[P + c], uy = mom[2 * P + c];
    const T ics2 = K.ics2, ics4 = K.ics4;
    const T u2 = ux * ux + uy * uy;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      T v = C::dev ? deq_k<T>(k, rho - (T)1, rho, ux, uy, u2, K.w[k], ics2, ics4)
                   : feq_k<T>(k, rho, ux, uy, u2, K.w[k], ics2, ics4);
      auto r = C::enc(v, k, K);
      fa[pop_index<B>(L, i, j, k)] = r;
      fb[pop_index<B>(L, i, j, k)] = r;
    }
  } else {
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      auto z = enc_zero<S>(k, K);
      fa[pop_index<B>(L, i, j, k)] = z;
      fb[pop_index<B>(L, i, j, k)] = z;
    }
  }
}

// rho/u of the stored populations on every fluid cell.
template <int S>
__global__ __launch_bounds__(256) void k_moments(const d2q9_lattice L,
                                                 const typename Codec<S>::raw* f,
                                                 typename Codec<S>::T* mom) {
  using C = Codec<S>;
  using T = typename C::T;
  const int64_t n = (int64_t)L.nx * L.ny;
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const int i = 1 + (int)(e / L.ny), j = 1 + (int)(e % L.ny);
  const int64_t c = cell_index(L, i, j);
  if (!(L.flags[c] & D2Q9_CELL_FLUID)) return;
  const int64_t P = L.plane;
  const Consts<T> K = load_consts<T>(L);
  T g[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) g[k] = ld<S>(L, K, f, i, j, k);
  T ffx = (T)0, ffy = (T)0;
  if (L.force_source == D2Q9_FSRC_FIELD) {
    const T* ff = static_cast<const T*>(L.force_field);
    ffx = ff[c];
    ffy = ff[P + c];
  }
  T drho, rho, ux, uy, fx, fy, px, py;
  if (L.forcing == D2Q9_FORCING_NONE)
    moments<T, D2Q9_FORCING_NONE, C::dev>(g, L, K, ffx, ffy, drho, rho, ux, uy, fx, fy, px, py);
  else
    moments<T, D2Q9_GUO_LINEAR, C::dev>(g, L, K, ffx, ffy, drho, rho, ux, uy, fx, fy, px, py);
  mom[c] = rho;
  mom[P + c] = ux;
  mom[2 * P + c] = uy;
}

template <typename T>
__global__ __launch_bounds__(RED_THREADS) void k_residue_partial(
    const d2q9_lattice L, const T* mom, T* mom_old, double* partial) {
  __shared__ double sh[RED_THREADS / 64];
  const int64_t n = (int64_t)L.nx * L.ny;
  double acc[6] = {0, 0, 0, 0, 0, 0};
  for (int64_t e = (int64_t)blockIdx.x * RED_THREADS + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * RED_THREADS) {
    const int i = 1 + (int)(e / L.ny), j = 1 + (int)(e % L.ny);
    const int64_t c = cell_index(L, i, j);
    if (!(L.flags[c] & D2Q9_CELL_FLUID)) continue;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      double v = (double)mom[q * L.plane + c];
      double o = (double)mom_old[q * L.plane + c];
      double d = v - o;
      acc[2 * q] += d * d;
      acc[2 * q + 1] += o * o;
      mom_old[q * L.plane + c] = mom[q * L.plane + c];
    }
  }
  for (int q = 0; q < 6; ++q) {
    double r = block_sum(acc[q], sh);
    if (threadIdx.x == 0) partial[q * RED_MAX_BLOCKS + blockIdx.x] = r;
  }
}

__global__ __launch_bounds__(256) void k_copy16(const uint4* __restrict__ src,
                                                uint4* __restrict__ dst, int64_t n) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const uint4 v = src[e];
  __builtin_nontemporal_store(v.x, &dst[e].x);
  __builtin_nontemporal_store(v.y, &dst[e].y);
  __builtin_nontemporal_store(v.z, &dst[e].z);
  __builtin_nontemporal_store(v.w, &dst[e].w);
}

int validate(const d2q9_lattice* L) {
  if (!L) return fail(D2Q9_E_ARG, "null lattice descriptor");
  if (L->nx < 1 || L->ny < 1) return fail(D2Q9_E_ARG, "nx and ny must be >= 1");
  if (L->pitch < L->ny + 2) return fail(D2Q9_E_ARG, "pitch < ny + 2");
  if (L->origin < 0 || L->plane < (int64_t)L->origin + (int64_t)(L->nx + 2) * L->pitch)
    return fail(D2Q9_E_ARG, "plane too small for origin + (nx + 2) * pitch");
  if (L->tile_i < 1 || L->tile_i > 4) return fail(D2Q9_E_ARG, "tile_i must be 1..4");
  if (L->storage < 0 || L->storage > 3) return fail(D2Q9_E_ARG, "bad storage");
  {
    const int cpl = L->storage == D2Q9_STORE_F64 ? 1 : (L->storage == D2Q9_STORE_F32 ? 2 : CPL16);
    const int B = 64 * cpl;
    if (L->cpl != cpl) return fail(D2Q9_E_ARG, "cpl does not match the storage");
    if (L->wall_masks && !L->wall_index)
      return fail(D2Q9_E_ARG, "wall_masks needs wall_index");
    if (L->storage == D2Q9_STORE_F32 || L->storage == D2Q9_STORE_BF16DEV)
      for (int k = 0; k < 9; ++k)
        if (L->dev_scale[k] != 1.0 || L->cf.scale[k] != 1.0f)
          return fail(D2Q9_E_ARG, "FP32 and BF16 storage keep dev_scale 1");
    if (L->tile_j != B) return fail(D2Q9_E_ARG, "tile_j must be 64 * cpl");
    if (L->nbj < (L->ny + B) / B + 1) return fail(D2Q9_E_ARG, "nbj too small");
    if (L->pop_len < (int64_t)(L->nx + 2) * L->nbj * 9 * B)
      return fail(D2Q9_E_ARG, "pop_len too small");
  }
  if (L->storage < 0 || L->storage > 3) return fail(D2Q9_E_ARG, "bad storage");
  if (L->collision < 0 || L->collision > 1) return fail(D2Q9_E_ARG, "bad collision");
  if (L->collision == D2Q9_MRT)
    for (int k = 0; k < 7; ++k)
      if (L->rates[k] != 1.0 || L->cf.rates[k] != 1.0f)
        return fail(D2Q9_E_UNSUPPORTED,
                    "MRT relaxes rows 0..6 at rate 1 (the reference's S); only rates 7 and 8 are free");
  if (L->forcing < 0 || L->forcing > 2) return fail(D2Q9_E_ARG, "bad forcing");
  if (L->force_source == D2Q9_FSRC_FIELD && !L->force_field)
    return fail(D2Q9_E_ARG, "force_source FIELD needs force_field");
  if (L->force_source == D2Q9_FSRC_FIELD && L->forcing == D2Q9_FORCING_NONE)
    return fail(D2Q9_E_UNSUPPORTED, "a force field needs a Guo forcing");
  if (!L->flags) return fail(D2Q9_E_ARG, "null flags");
  if ((L->bc_nodes == nullptr) != (L->bc_slot == nullptr))
    return fail(D2Q9_E_ARG, "bc_nodes and bc_slot go together");
  return 0;
}

// One auxiliary stream and a fork/join event pair per caller stream (and
// device), created on first use and kept for the life of the process
// (default priority).  Keying them on the caller's stream means that two
// host threads driving lattices on different streams never share an event:
// each call's fork and join order only its own stream's work, and a graph
// capture on one stream never pulls in another thread's eager work.
// `order` makes the fork ... join sequence of one call atomic with respect
// to other calls on the same caller stream.
struct Aux {
  hipStream_t stream = nullptr;
  hipEvent_t fork = nullptr, join = nullptr;
  std::mutex order;
};
constexpr size_t MAX_AUX = 4096;
std::map<std::pair<int, hipStream_t>, std::unique_ptr<Aux>> g_aux;
std::mutex g_aux_mutex;

int aux_for_stream(hipStream_t s, Aux** out) {
  int dev = 0;
  if (int r = check_hip(hipGetDevice(&dev), "hipGetDevice")) return r;
  std::lock_guard<std::mutex> lock(g_aux_mutex);
  std::unique_ptr<Aux>& slot = g_aux[std::make_pair(dev, s)];
  if (!slot) {
    if (g_aux.size() > MAX_AUX) {
      g_aux.erase(std::make_pair(dev, s));
      return fail(D2Q9_E_TOO_LARGE, "more than 4096 distinct caller streams");
    }
    std::unique_ptr<Aux> a(new Aux);
    if (int r = check_hip(hipStreamCreateWithFlags(&a->stream, hipStreamNonBlocking),
                          "hipStreamCreateWithFlags"))
      return r;
    if (int r = check_hip(hipEventCreateWithFlags(&a->fork, hipEventDisableTiming),
                          "hipEventCreateWithFlags"))
      return r;
    if (int r = check_hip(hipEventCreateWithFlags(&a->join, hipEventDisableTiming),
                          "hipEventCreateWithFlags"))
      return r;
    slot = std::move(a);
  }
  *out = slot.get();
  return 0;
}

// fork == false: the general/fix-up kernel runs on `s` too, before the
// bulk kernel (for callers that already run on the auxiliary stream).
template <int S, int MODEL, int FORCING, int PHASE>
int launch_sweep(const d2q9_lattice* L, const void* fsrc, void* fdst,
                 const void* msrc, void* mdst, int write_mom, hipStream_t s,
                 bool fork = true) {
  using C = Codec<S>;
  const int ntj = (L->ny + L->tile_j - 1) / L->tile_j;
  const int nti = (L->nx + L->tile_i - 1) / L->tile_i;
  const int64_t ntiles = (int64_t)ntj * nti;
  if (ntiles > 0x7fffffff) return fail(D2Q9_E_ARG, "lattice too large");
  dim3 block(64, L->tile_i);
  auto fs = static_cast<const typename C::raw*>(fsrc);
  auto fd = static_cast<typename C::raw*>(fdst);
  auto ms = static_cast<const typename C::T*>(msrc);
  auto md = static_cast<typename C::T*>(mdst);
  if (L->tiles == nullptr) {
    hipLaunchKernelGGL((k_general_fix<S, MODEL, FORCING, PHASE>), dim3((unsigned)ntiles),
                       block, 0, s, *L, fs, fd, ms, md, write_mom, ntj, (int)ntiles, 0);
    return check_hip(hipGetLastError(), "k_general_fix launch");
  }
  if ((int64_t)L->n_bulk + L->n_general > ntiles)
    return fail(D2Q9_E_ARG, "n_bulk + n_general exceeds the tile count");
  if (L->n_fix > 0 && !L->fix_cells) return fail(D2Q9_E_ARG, "null fix_cells");
  if (L->n_bulk > 0 && !L->bulk_masks) return fail(D2Q9_E_ARG, "null bulk_masks");
  const int wg = 64 * L->tile_i;
  const int nfixb = (L->n_fix + wg - 1) / wg;
  // Fused steps of the float forms run as one kernel (measured on the
  // 8192x2048 cylinder: FP32 227 -> 221 us, BF16 119 -> 112 us, FP16
  // 122 -> 120 us).  FP64 keeps the two overlapped kernels: its bulk sweep
  // is HBM-bound at full occupancy and the one-kernel form measured 7 us
  // slower.  The general path is compiled into the one kernel only for FP32
  // (for the 16-bit forms it would cost occupancy).
  if constexpr (PHASE == PHASE_FUSED && S != D2Q9_STORE_F64) {
    if (L->n_bulk > 0 && L->n_general == 0) {
      const unsigned nb = (unsigned)(nfixb + L->n_bulk);
      hipLaunchKernelGGL((k_step_all<S, MODEL, FORCING, PHASE, false>), dim3(nb), block, 0, s,
                         *L, fs, fd, ms, md, write_mom, ntj, nfixb);
      return check_hip(hipGetLastError(), "k_step_all launch");
    }
  }
  if constexpr (S == D2Q9_STORE_F32 && PHASE == PHASE_FUSED) {
    if (L->n_bulk > 0) {
      const unsigned nb = (unsigned)(L->n_general + nfixb + L->n_bulk);
      hipLaunchKernelGGL((k_step_all<S, MODEL, FORCING, PHASE, true>), dim3(nb), block, 0, s,
                         *L, fs, fd, ms, md, write_mom, ntj, nfixb);
      return check_hip(hipGetLastError(), "k_step_all launch");
    }
  }
  hipStream_t aux = s;
  Aux* a = nullptr;
  std::unique_lock<std::mutex> order;
  if (fork && L->n_bulk > 0 && L->n_general + nfixb > 0) {
    if (int r = aux_for_stream(s, &a)) return r;
    order = std::unique_lock<std::mutex>(a->order);
    aux = a->stream;
    if (int r = check_hip(hipEventRecord(a->fork, s), "hipEventRecord fork")) return r;
    if (int r = check_hip(hipStreamWaitEvent(aux, a->fork, 0), "hipStreamWaitEvent fork"))
      return r;
  }
  if (L->n_general + nfixb > 0) {
    hipLaunchKernelGGL((k_general_fix<S, MODEL, FORCING, PHASE>),
                       dim3((unsigned)(L->n_general + nfixb)), block, 0, aux, *L, fs, fd, ms,
                       md, write_mom, ntj, L->n_general, L->n_bulk);
    if (int r = check_hip(hipGetLastError(), "k_general_fix launch")) return r;
  }
  if (L->n_bulk > 0) {
    hipLaunchKernelGGL((k_bulk<S, MODEL, FORCING, PHASE>), dim3((unsigned)L->n_bulk), block,
                       0, s, *L, fs, fd, md, write_mom, ntj);
    if (int r = check_hip(hipGetLastError(), "k_bulk launch")) return r;
  }
  if (a) {
    if (int r = check_hip(hipEventRecord(a->join, aux), "hipEventRecord join")) return r;
    if (int r = check_hip(hipStreamWaitEvent(s, a->join, 0), "hipStreamWaitEvent join"))
      return r;
  }
  return 0;
}

template <int S, int PHASE>
int dispatch_model(const d2q9_lattice* L, const void* a, void* b, const void* c,
                   void* d, int wm, hipStream_t s, bool fork = true) {
  const int key = L->collision * 3 + L->forcing;
  switch (key) {
    case 0: return launch_sweep<S, D2Q9_BGK, D2Q9_FORCING_NONE, PHASE>(L, a, b, c, d, wm, s, fork);
    case 1: return launch_sweep<S, D2Q9_BGK, D2Q9_GUO_LINEAR, PHASE>(L, a, b, c, d, wm, s, fork);
    case 2: return launch_sweep<S, D2Q9_BGK, D2Q9_GUO_SECOND_ORDER, PHASE>(L, a, b, c, d, wm, s, fork);
    case 3: return launch_sweep<S, D2Q9_MRT, D2Q9_FORCING_NONE, PHASE>(L, a, b, c, d, wm, s, fork);
    case 4: return launch_sweep<S, D2Q9_MRT, D2Q9_GUO_LINEAR, PHASE>(L, a, b, c, d, wm, s, fork);
    case 5: return launch_sweep<S, D2Q9_MRT, D2Q9_GUO_SECOND_ORDER, PHASE>(L, a, b, c, d, wm, s, fork);
  }
  return fail(D2Q9_E_ARG, "bad collision/forcing");
}

template <int PHASE>
int dispatch(const d2q9_lattice* L, const void* fsrc, void* fdst, const void* msrc,
             void* mdst, int wm, void* stream) {
  if (int r = validate(L)) return r;
  if (!fsrc || !fdst || !msrc) return fail(D2Q9_E_ARG, "null buffer");
  if (PHASE != PHASE_STREAM && !mdst) return fail(D2Q9_E_ARG, "null mom_dst");
  if (fsrc == fdst) return fail(D2Q9_E_ARG, "f_src and f_dst must differ");
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (PHASE == PHASE_STREAM) {
    // streaming never touches the collision model: one instantiation each
    switch (L->storage) {
      case D2Q9_STORE_F64: return launch_sweep<D2Q9_STORE_F64, D2Q9_BGK, D2Q9_FORCING_NONE, PHASE>(L, fsrc, fdst, msrc, mdst, 0, s);
      case D2Q9_STORE_F32: return launch_sweep<D2Q9_STORE_F32, D2Q9_BGK, D2Q9_FORCING_NONE, PHASE>(L, fsrc, fdst, msrc, mdst, 0, s);
      case D2Q9_STORE_F16DEV: return launch_sweep<D2Q9_STORE_F16DEV, D2Q9_BGK, D2Q9_FORCING_NONE, PHASE>(L, fsrc, fdst, msrc, mdst, 0, s);
      case D2Q9_STORE_BF16DEV: return launch_sweep<D2Q9_STORE_BF16DEV, D2Q9_BGK, D2Q9_FORCING_NONE, PHASE>(L, fsrc, fdst, msrc, mdst, 0, s);
    }
  } else {
    switch (L->storage) {
      case D2Q9_STORE_F64: return dispatch_model<D2Q9_STORE_F64, PHASE>(L, fsrc, fdst, msrc, mdst, wm, s);
      case D2Q9_STORE_F32: return dispatch_model<D2Q9_STORE_F32, PHASE>(L, fsrc, fdst, msrc, mdst, wm, s);
      case D2Q9_STORE_F16DEV: return dispatch_model<D2Q9_STORE_F16DEV, PHASE>(L, fsrc, fdst, msrc, mdst, wm, s);
      case D2Q9_STORE_BF16DEV: return dispatch_model<D2Q9_STORE_BF16DEV, PHASE>(L, fsrc, fdst, msrc, mdst, wm, s);
    }
  }
  return fail(D2Q9_E_ARG, "bad storage");
}

}  // namespace

int d2q9::validate_lattice(const d2q9_lattice* L) { return validate(L); }

extern "C" {


int d2q9_init_eq(const d2q9_lattice* L, void* f_a, void* f_b, const void* mom,
                 void* stream) {
  if (int r = validate(L)) return r;
  if (!f_a || !f_b || !mom) return fail(D2Q9_E_ARG, "null buffer");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int64_t n = (int64_t)(L->nx + 2) * (L->ny + 2);
  const unsigned blocks = (unsigned)((n + 255) / 256);
  switch (L->storage) {
#define D2Q9_INIT(SV)                                                          \
  case SV:                                                                     \
    hipLaunchKernelGGL((k_init<SV>), dim3(blocks), dim3(256), 0, s, *L,        \
                       static_cast<Codec<SV>::raw*>(f_a),                      \
                       static_cast<Codec<SV>::raw*>(f_b),                      \
                       static_cast<const Codec<SV>::T*>(mom));                 \
    break;
    D2Q9_INIT(D2Q9_STORE_F64)
    D2Q9_INIT(D2Q9_STORE_F32)
    D2Q9_INIT(D2Q9_STORE_F16DEV)
    D2Q9_INIT(D2Q9_STORE_BF16DEV)
#undef D2Q9_INIT
    default: return fail(D2Q9_E_ARG, "bad storage");
  }
  return check_hip(hipGetLastError(), "k_init launch");
}

int d2q9_collide(const d2q9_lattice* L, const void* f_src, void* f_dst,
                 const void* mom_src, void* mom_dst, int write_moments,
                 void* stream) {
  return dispatch<PHASE_COLLIDE>(L, f_src, f_dst, mom_src, mom_dst, write_moments, stream);
}

int d2q9_stream_bc(const d2q9_lattice* L, const void* f_src, void* f_dst,
                   const void* mom_src, void* stream) {
  return dispatch<PHASE_STREAM>(L, f_src, f_dst, mom_src, nullptr, 0, stream);
}

int d2q9_fused_step(const d2q9_lattice* L, const void* f_src, void* f_dst,
                    const void* mom_src, void* mom_dst, int write_moments,
                    void* stream) {
  if (mom_src == mom_dst) return fail(D2Q9_E_ARG, "mom_src and mom_dst must differ");
  return dispatch<PHASE_FUSED>(L, f_src, f_dst, mom_src, mom_dst, write_moments, stream);
}

int d2q9_fused_step2(const d2q9_lattice* L, const d2q9_lattice* L_rim_a,
                     const d2q9_lattice* L_rim_b, const int32_t* deep_tiles,
                     const uint8_t* deep_walls, int32_t n_deep, const void* f_src,
                     void* f_dst, void* f_tmp,
                     const void* mom_src, void* mom_dst, void* mom_tmp,
                     int write_moments, void* stream) {
  for (const d2q9_lattice* X : {L, L_rim_a, L_rim_b})
    if (int r = validate(X)) return r;
  for (const d2q9_lattice* X : {L_rim_a, L_rim_b}) {
    if (X->storage != L->storage || X->nx != L->nx || X->ny != L->ny ||
        X->pitch != L->pitch || X->plane != L->plane || X->origin != L->origin ||
        X->nbj != L->nbj || X->tile_i != L->tile_i || X->collision != L->collision ||
        X->forcing != L->forcing || !X->tiles)
      return fail(D2Q9_E_ARG, "rim descriptors must describe the same lattice, with tile lists");
  }
  if (n_deep < 0 || (n_deep > 0 && !deep_tiles)) return fail(D2Q9_E_ARG, "bad deep tile list");
  if (!f_src || !f_dst || !f_tmp || !mom_src || !mom_dst || !mom_tmp)
    return fail(D2Q9_E_ARG, "null buffer");
  if (f_src == f_dst || f_src == f_tmp || f_dst == f_tmp || mom_src == mom_dst ||
      mom_src == mom_tmp || mom_dst == mom_tmp)
    return fail(D2Q9_E_ARG, "f_src, f_dst, f_tmp (and the moments) must all differ");
  const int ntj = (L->ny + L->tile_j - 1) / L->tile_j;
  if ((int64_t)n_deep > (int64_t)ntj * ((L->nx + DEEP_ROWS - 1) / DEEP_ROWS))
    return fail(D2Q9_E_ARG, "more deep tiles than the lattice holds");
  hipStream_t s = static_cast<hipStream_t>(stream);
  // with no rim (every cell in a deep tile: a channel periodic in x with
  // straight walls at y) the step is the deep kernel alone on the caller's
  // stream, with no fork and join
  const bool rim_empty = L_rim_a->n_bulk + L_rim_a->n_general + L_rim_a->n_fix == 0 &&
                         L_rim_b->n_bulk + L_rim_b->n_general + L_rim_b->n_fix == 0;
  Aux* a = nullptr;
  std::unique_lock<std::mutex> order;
  if (!rim_empty) {
    if (int r = aux_for_stream(s, &a)) return r;
    order = std::unique_lock<std::mutex>(a->order);
    if (int r = check_hip(hipEventRecord(a->fork, s), "hipEventRecord fork")) return r;
  }
  if (n_deep > 0) {
    const dim3 grid((unsigned)n_deep), block(64, DEEP_WAVES);
    const int key = L->collision * 3 + L->forcing;
#define D2Q9_DEEP_XW(SV, M, F, XW)                                                       \
  hipLaunchKernelGGL((k_deep2<SV, M, F, XW>), grid, block, 0, s, *L, deep_tiles, deep_walls, \
                     n_deep, ntj, static_cast<const Codec<SV>::raw*>(f_src),             \
                     static_cast<Codec<SV>::raw*>(f_dst),                                \
                     static_cast<Codec<SV>::T*>(mom_dst), write_moments)
#define D2Q9_DEEP(SV, KEY, M, F)                  \
  case KEY:                                       \
    if (L->x_periodic)                            \
      D2Q9_DEEP_XW(SV, M, F, true);               \
    else                                          \
      D2Q9_DEEP_XW(SV, M, F, false);              \
    break;
#define D2Q9_DEEP_ALL(SV)                                          \
  switch (key) {                                                   \
    D2Q9_DEEP(SV, 0, D2Q9_BGK, D2Q9_FORCING_NONE)                  \
    D2Q9_DEEP(SV, 1, D2Q9_BGK, D2Q9_GUO_LINEAR)                    \
    D2Q9_DEEP(SV, 2, D2Q9_BGK, D2Q9_GUO_SECOND_ORDER)              \
    D2Q9_DEEP(SV, 3, D2Q9_MRT, D2Q9_FORCING_NONE)                  \
    D2Q9_DEEP(SV, 4, D2Q9_MRT, D2Q9_GUO_LINEAR)                    \
    D2Q9_DEEP(SV, 5, D2Q9_MRT, D2Q9_GUO_SECOND_ORDER)              \
    default: return fail(D2Q9_E_ARG, "bad collision/forcing");     \
  }
    if (L->storage == D2Q9_STORE_F64) {
      D2Q9_DEEP_ALL(D2Q9_STORE_F64)
    } else if (L->storage == D2Q9_STORE_F32) {
      D2Q9_DEEP_ALL(D2Q9_STORE_F32)
    } else if (L->storage == D2Q9_STORE_F16DEV) {
      D2Q9_DEEP_ALL(D2Q9_STORE_F16DEV)
    } else {
      D2Q9_DEEP_ALL(D2Q9_STORE_BF16DEV)
    }
#undef D2Q9_DEEP_ALL
#undef D2Q9_DEEP
#undef D2Q9_DEEP_XW
    if (int r = check_hip(hipGetLastError(), "k_deep2 launch")) return r;
  }
  if (rim_empty) return 0;
  if (int r = check_hip(hipStreamWaitEvent(a->stream, a->fork, 0), "hipStreamWaitEvent fork"))
    return r;
  // rim: step n + 1 of the rim tiles and the cells they pull from into
  // f_tmp, then step n + 2 of the rim tiles into f_dst, on the aux stream
  // (enqueued after the deep tiles, whose workgroups then reach the CUs
  // first)
  auto rim = [&](const d2q9_lattice* X, const void* fs, void* fd, const void* ms, void* md,
                 int wm) {
    hipStream_t as = a->stream;
    switch (L->storage) {
      case D2Q9_STORE_F64:
        return dispatch_model<D2Q9_STORE_F64, PHASE_FUSED>(X, fs, fd, ms, md, wm, as, false);
      case D2Q9_STORE_F32:
        return dispatch_model<D2Q9_STORE_F32, PHASE_FUSED>(X, fs, fd, ms, md, wm, as, false);
      case D2Q9_STORE_F16DEV:
        return dispatch_model<D2Q9_STORE_F16DEV, PHASE_FUSED>(X, fs, fd, ms, md, wm, as, false);
      default:
        return dispatch_model<D2Q9_STORE_BF16DEV, PHASE_FUSED>(X, fs, fd, ms, md, wm, as, false);
    }
  };
  if (int r = rim(L_rim_a, f_src, f_tmp, mom_src, mom_tmp, 0)) return r;
  if (int r = rim(L_rim_b, f_tmp, f_dst, mom_tmp, mom_dst, write_moments)) return r;
  if (int r = check_hip(hipEventRecord(a->join, a->stream), "hipEventRecord join")) return r;
  return check_hip(hipStreamWaitEvent(s, a->join, 0), "hipStreamWaitEvent join");
}

int d2q9_deep_rows(void) { return DEEP_ROWS; }

int d2q9_moments(const d2q9_lattice* L, const void* f, void* mom, void* stream) {
  if (int r = validate(L)) return r;
  if (!f || !mom) return fail(D2Q9_E_ARG, "null buffer");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int64_t n = (int64_t)L->nx * L->ny;
  const unsigned blocks = (unsigned)((n + 255) / 256);
  switch (L->storage) {
#define D2Q9_MOM(SV)                                                           \
  case SV:                                                                     \
    hipLaunchKernelGGL((k_moments<SV>), dim3(blocks), dim3(256), 0, s, *L,     \
                       static_cast<const Codec<SV>::raw*>(f),                  \
                       static_cast<Codec<SV>::T*>(mom));                       \
    break;
    D2Q9_MOM(D2Q9_STORE_F64)
    D2Q9_MOM(D2Q9_STORE_F32)
    D2Q9_MOM(D2Q9_STORE_F16DEV)
    D2Q9_MOM(D2Q9_STORE_BF16DEV)
#undef D2Q9_MOM
    default: return fail(D2Q9_E_ARG, "bad storage");
  }
  return check_hip(hipGetLastError(), "k_moments launch");
}

int64_t d2q9_reduce_work_len(const d2q9_lattice* L) {
  (void)L;
  return (int64_t)6 * RED_MAX_BLOCKS;
}

int d2q9_mass_fp64(const d2q9_lattice* L, const void* f, double* work,
                   double* out, void* stream) {
  if (int r = validate(L)) return r;
  if (!f || !work || !out) return fail(D2Q9_E_ARG, "null buffer");
  return launch_mass(L, f, work, out, static_cast<hipStream_t>(stream));
}

int d2q9_residues(const d2q9_lattice* L, const void* mom, void* mom_old,
                  double* work, double* out, void* stream) {
  if (int r = validate(L)) return r;
  if (!mom || !mom_old || !work || !out) return fail(D2Q9_E_ARG, "null buffer");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int nb = red_blocks(L);
  if (L->storage == D2Q9_STORE_F64) {
    hipLaunchKernelGGL((k_residue_partial<double>), dim3(nb), dim3(RED_THREADS), 0, s,
                       *L, static_cast<const double*>(mom), static_cast<double*>(mom_old), work);
  } else {
    hipLaunchKernelGGL((k_residue_partial<float>), dim3(nb), dim3(RED_THREADS), 0, s,
                       *L, static_cast<const float*>(mom), static_cast<float*>(mom_old), work);
  }
  hipLaunchKernelGGL((k_finish<6>), dim3(1), dim3(RED_THREADS), 0, s, work, nb, out);
  return check_hip(hipGetLastError(), "k_residue launch");
}

int d2q9_copy(void* dst, const void* src, int64_t bytes, void* stream) {
  if (!dst || !src || bytes < 0 || (bytes % 16) != 0)
    return fail(D2Q9_E_ARG, "copy needs non-null buffers and a multiple of 16 bytes");
  const int64_t n = bytes / 16;
  int64_t blocks = (n + 255) / 256;
  if (blocks > 0x7fffffff) return fail(D2Q9_E_TOO_LARGE, "copy too large");
  if (blocks < 1) return 0;
  hipLaunchKernelGGL(k_copy16, dim3((unsigned)blocks), dim3(256), 0,
                     static_cast<hipStream_t>(stream), static_cast<const uint4*>(src),
                     static_cast<uint4*>(dst), n);
  return check_hip(hipGetLastError(), "k_copy16 launch");
}

int d2q9_graph_begin(void* stream) {
  return check_hip(hipStreamBeginCapture(static_cast<hipStream_t>(stream),
                                         hipStreamCaptureModeThreadLocal),
                   "hipStreamBeginCapture");
}

int d2q9_graph_end(void* stream, void** graph_exec) {
  if (!graph_exec) return fail(D2Q9_E_ARG, "null graph_exec");
  hipGraph_t g = nullptr;
  if (int r = check_hip(hipStreamEndCapture(static_cast<hipStream_t>(stream), &g),
                        "hipStreamEndCapture"))
    return r;
  hipGraphExec_t ge = nullptr;
  int r = check_hip(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0), "hipGraphInstantiate");
  (void)hipGraphDestroy(g);
  if (r) return r;
  *graph_exec = ge;
  return 0;
}

int d2q9_graph_launch(void* graph_exec, void* stream) {
  if (!graph_exec) return fail(D2Q9_E_ARG, "null graph_exec");
  return check_hip(hipGraphLaunch(static_cast<hipGraphExec_t>(graph_exec),
                                  static_cast<hipStream_t>(stream)),
                   "hipGraphLaunch");
}

int d2q9_graph_destroy(void* graph_exec) {
  if (!graph_exec) return 0;
  return check_hip(hipGraphExecDestroy(static_cast<hipGraphExec_t>(graph_exec)),
                   "hipGraphExecDestroy");
}

const char* d2q9_last_error(void) { return g_last_error.c_str(); }

int d2q9_abi_layout(int64_t* out, int n) {
  const int64_t v[] = {
      (int64_t)sizeof(d2q9_lattice), (int64_t)sizeof(d2q9_bc_elem),
      (int64_t)sizeof(d2q9_bc_node), (int64_t)offsetof(d2q9_lattice, plane),
      (int64_t)offsetof(d2q9_lattice, cpl), (int64_t)offsetof(d2q9_lattice, pop_len),
      (int64_t)offsetof(d2q9_lattice, storage), (int64_t)offsetof(d2q9_lattice, tile_j),
      (int64_t)offsetof(d2q9_lattice, omega), (int64_t)offsetof(d2q9_lattice, w),
      (int64_t)offsetof(d2q9_lattice, gravity), (int64_t)offsetof(d2q9_lattice, rates),
      (int64_t)offsetof(d2q9_lattice, dev_scale), (int64_t)offsetof(d2q9_lattice, cf),
      (int64_t)offsetof(d2q9_lattice, flags), (int64_t)offsetof(d2q9_lattice, tiles),
      (int64_t)offsetof(d2q9_lattice, n_bulk), (int64_t)offsetof(d2q9_lattice, fix_cells),
      (int64_t)offsetof(d2q9_lattice, n_fix), (int64_t)offsetof(d2q9_lattice, bulk_masks),
      (int64_t)offsetof(d2q9_lattice, bc_slot),
      (int64_t)offsetof(d2q9_lattice, bc_nodes), (int64_t)offsetof(d2q9_lattice, bc_elems),
      (int64_t)offsetof(d2q9_lattice, force_field),
      (int64_t)offsetof(d2q9_lattice, wall_index), (int64_t)offsetof(d2q9_lattice, wall_masks),
      (int64_t)sizeof(d2q9_batch_entry), (int64_t)offsetof(d2q9_batch_entry, f),
      (int64_t)offsetof(d2q9_batch_entry, mom), (int64_t)sizeof(d2q9_batch_item)};
  const int m = (int)(sizeof(v) / sizeof(v[0]));
  if (!out || n < m) return fail(D2Q9_E_ARG, "abi layout needs 30 slots");
  for (int q = 0; q < m; ++q) out[q] = v[q];
  return m;
}

}  // extern "C"
