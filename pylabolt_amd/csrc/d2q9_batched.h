// Batched persistent sweep: many lattices of different sizes, boundaries and
// collision models advance one fused time step in a single launch.
//
// The work of a step is a list of items -- bulk tiles, general tiles and
// chunks of fix-up cells of every lattice -- that write disjoint cells, so
// a fixed grid of workgroups can take them in any order.  The list is cut
// into eight queues, one per XCD (workgroup b runs on XCD b % 8), each
// holding whole lattices, so that the tiles that share cache lines stay in
// one L2.  Policy 0 walks each queue statically; policy 1 claims items with
// an atomic counter per queue and, once its own queue is empty, helps the
// others, which evens out lattices of uneven cost; a claim takes `chunk`
// consecutive items.  The counters of parity
// p are used by the step of parity p; that step clears the other half for
// the next step, so a HIP graph can replay the steps with no reset node.
#pragma once
#include <hip/hip_runtime.h>

#include "d2q9_sweep.h"

namespace d2q9 {
namespace batched {

constexpr int PHASE = PHASE_FUSED;

template <int S, int MODEL, int FORCING, bool GEN>
__device__ __forceinline__ void run_item(const d2q9_batch_entry& E, int parity, int code,
                                         int write_mom) {
  using C = Codec<S>;
  const d2q9_lattice& L = E.L;
  auto fs = static_cast<const typename C::raw*>(E.f[parity]);
  auto fd = static_cast<typename C::raw*>(E.f[1 - parity]);
  auto ms = static_cast<const typename C::T*>(E.mom[parity]);
  auto md = static_cast<typename C::T*>(E.mom[1 - parity]);
  const int ntj = (L.ny + L.tile_j - 1) / L.tile_j;
  if (code < 0) {
    fix_chunk<S, MODEL, FORCING, PHASE>(L, fs, fd, ms, md, write_mom,
                                        (-code - 1) * 64 * (int)blockDim.y);
  } else if (code < L.n_bulk) {
    bulk_tile<S, MODEL, FORCING, PHASE>(L, fs, fd, md, write_mom, ntj, code);
  } else {
    // a general item in a batch declared without them is skipped
    if constexpr (GEN)
      general_tile<S, MODEL, FORCING, PHASE>(L, fs, fd, ms, md, write_mom, ntj, L.tiles[code]);
  }
}

template <int S, bool GEN>
__device__ __forceinline__ void dispatch_item(const d2q9_batch_entry* __restrict__ entries,
                                              const d2q9_batch_item it, int parity,
                                              int write_mom) {
  const d2q9_batch_entry& E = entries[it.lattice];
  switch (E.L.collision * 3 + E.L.forcing) {
    case 0: run_item<S, D2Q9_BGK, D2Q9_FORCING_NONE, GEN>(E, parity, it.code, write_mom); break;
    case 1: run_item<S, D2Q9_BGK, D2Q9_GUO_LINEAR, GEN>(E, parity, it.code, write_mom); break;
    case 2: run_item<S, D2Q9_BGK, D2Q9_GUO_SECOND_ORDER, GEN>(E, parity, it.code, write_mom); break;
    case 3: run_item<S, D2Q9_MRT, D2Q9_FORCING_NONE, GEN>(E, parity, it.code, write_mom); break;
    case 4: run_item<S, D2Q9_MRT, D2Q9_GUO_LINEAR, GEN>(E, parity, it.code, write_mom); break;
    default: run_item<S, D2Q9_MRT, D2Q9_GUO_SECOND_ORDER, GEN>(E, parity, it.code, write_mom); break;
  }
}

// W: waves per SIMD the register allocation must allow (1: no limit
// beyond the compiler's own choice).
// GEN false: no lattice of the batch has general tiles (solid cells), and
// the general-tile path is left out of the kernel (fewer registers).
template <int S, int W, bool GEN>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(W))) void k_batched(const d2q9_batch_entry* __restrict__ entries,
                                                 const d2q9_batch_item* __restrict__ items,
                                                 const int32_t* __restrict__ qb,
                                                 int32_t* __restrict__ counters, int parity,
                                                 int policy, int chunk, int write_mom) {
  __shared__ int s_claim;
  const int xcd = (int)blockIdx.x & 7;
  const int r = (int)blockIdx.x >> 3;
  const int per = (int)gridDim.x >> 3;
  const int tid = (int)(threadIdx.y * blockDim.x + threadIdx.x);
  // clear the counters the next step (other parity) will use; per-lane
  // addresses keep this a vector store
  if (blockIdx.x == 0 && tid < 8) counters[(1 - parity) * 8 + tid] = 0;
  if (policy == 0) {
    for (int it = qb[xcd] + r; it < qb[xcd + 1]; it += per)
      dispatch_item<S, GEN>(entries, items[it], parity, write_mom);
    return;
  }
  for (int s = 0; s < 8; ++s) {
    const int q = (xcd + s) & 7;
    const int lo = qb[q], n = qb[q + 1] - lo;
    while (true) {
      if (tid == 0) s_claim = atomicAdd(&counters[parity * 8 + q], chunk);
      __syncthreads();
      const int k0 = s_claim;
      __syncthreads();
      if (k0 >= n) break;
      const int k1 = min(k0 + chunk, n);
      for (int k = k0; k < k1; ++k)
        dispatch_item<S, GEN>(entries, items[lo + k], parity, write_mom);
    }
  }
}


// Launches k_batched<S>; instantiated once per storage type (FP64, FP32)
// in its own translation unit so they build in parallel.
template <int S>
hipError_t launch(const d2q9_batch_entry* entries, const d2q9_batch_item* items,
                  const int32_t* qb, int32_t* counters, int parity, int tile_i,
                  int grid, int policy, int chunk, int waves, int has_general,
                  int write_mom, hipStream_t s) {
  const dim3 g(grid), b(64, tile_i);
#define D2Q9_BLAUNCH(WV, GN)                                                            \
  hipLaunchKernelGGL((k_batched<S, WV, GN>), g, b, 0, s, entries, items, qb, counters, \
                     parity, policy, chunk, write_mom)
  switch (waves * 2 + (has_general ? 1 : 0)) {
    case 0: D2Q9_BLAUNCH(1, false); break;
    case 1: D2Q9_BLAUNCH(1, true); break;
    case 12: D2Q9_BLAUNCH(6, false); break;
    case 13: D2Q9_BLAUNCH(6, true); break;
    case 16: D2Q9_BLAUNCH(8, false); break;
    case 17: D2Q9_BLAUNCH(8, true); break;
    default:
      return hipErrorInvalidValue;
  }
#undef D2Q9_BLAUNCH
  return hipGetLastError();
}

#define D2Q9_BATCHED_DECLARE(SV)                                                        \
  extern template hipError_t launch<SV>(const d2q9_batch_entry*, const d2q9_batch_item*, \
                                        const int32_t*, int32_t*, int, int, int, int,   \
                                        int, int, int, int, hipStream_t);
D2Q9_BATCHED_DECLARE(D2Q9_STORE_F64)
D2Q9_BATCHED_DECLARE(D2Q9_STORE_F32)
#undef D2Q9_BATCHED_DECLARE

}  // namespace batched
}  // namespace d2q9
