// FP64 reductions over the fluid cells of one lattice: a deterministic
// two-level tree (one partial sum per workgroup, then one workgroup adds the
// partials in block order), so a run is reproducible bit for bit.
//
// Used by d2q9_mass_fp64 and d2q9_residues (d2q9_step.hip) and by
// d2q9_mass_fp64_static (d2q9_noalloc.hip); the two mass entry points launch
// the same kernels and differ only in where the partial sums live.
#pragma once

#include <hip/hip_runtime.h>

#include "d2q9_device.h"

namespace d2q9 {

// The argument checks of every single-lattice entry point (d2q9_step.hip):
// 0, or D2Q9_E_ARG with the error text set.
int validate_lattice(const d2q9_lattice* L);

constexpr int RED_THREADS = 256;
constexpr int RED_MAX_BLOCKS = 2048;

__device__ __forceinline__ double block_sum(double v, double* sh) {
  // wave-level tree then across the 4 waves of the workgroup
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) sh[wave] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) {
    for (int w = 0; w < RED_THREADS / 64; ++w) r += sh[w];
  }
  __syncthreads();
  return r;
}

// partial[blockIdx.x] = sum of all populations of the fluid cells the
// workgroup visits (grid-stride over the interior cells).
template <int S>
__global__ __launch_bounds__(RED_THREADS) void k_mass_partial(
    const d2q9_lattice L, const typename Codec<S>::raw* f, double* partial) {
  __shared__ double sh[RED_THREADS / 64];
  const int64_t n = (int64_t)L.nx * L.ny;
  double acc = 0.0;
  for (int64_t e = (int64_t)blockIdx.x * RED_THREADS + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * RED_THREADS) {
    const int i = 1 + (int)(e / L.ny), j = 1 + (int)(e % L.ny);
    const int64_t c = cell_index(L, i, j);
    if (!(L.flags[c] & D2Q9_CELL_FLUID)) continue;
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < 9; ++k)
      s += Codec<S>::f64(f[pop_index<64 * Codec<S>::CPL>(L, i, j, k)], k, L);
    acc += s;
  }
  double r = block_sum(acc, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = r;
}

// out[q] = sum over b < nblocks of partial[q * RED_MAX_BLOCKS + b].
template <int NV>
__global__ __launch_bounds__(RED_THREADS) void k_finish(const double* partial,
                                                        int nblocks, double* out) {
  __shared__ double sh[RED_THREADS / 64];
  for (int q = 0; q < NV; ++q) {
    double acc = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += RED_THREADS) acc += partial[q * RED_MAX_BLOCKS + b];
    double r = block_sum(acc, sh);
    if (threadIdx.x == 0) out[q] = r;
  }
}

// Workgroups of a reduction over L's interior: about 8 cells per thread,
// at most RED_MAX_BLOCKS.
inline int red_blocks(const d2q9_lattice* L) {
  int64_t n = (int64_t)L->nx * L->ny;
  int64_t b = (n + RED_THREADS * 8 - 1) / (RED_THREADS * 8);
  if (b < 1) b = 1;
  if (b > RED_MAX_BLOCKS) b = RED_MAX_BLOCKS;
  return (int)b;
}

// The mass reduction into *out with red_blocks(L) partial sums in `work`
// (L already validated).  Returns 0, D2Q9_E_ARG or a hipError_t.
inline int launch_mass(const d2q9_lattice* L, const void* f, double* work, double* out,
                       hipStream_t s) {
  const int nb = red_blocks(L);
  switch (L->storage) {
#define D2Q9_MASS(SV)                                                          \
  case SV:                                                                     \
    hipLaunchKernelGGL((k_mass_partial<SV>), dim3(nb), dim3(RED_THREADS), 0,  \
                       s, *L, static_cast<const Codec<SV>::raw*>(f), work);    \
    break;
    D2Q9_MASS(D2Q9_STORE_F64)
    D2Q9_MASS(D2Q9_STORE_F32)
    D2Q9_MASS(D2Q9_STORE_F16DEV)
    D2Q9_MASS(D2Q9_STORE_BF16DEV)
#undef D2Q9_MASS
    default:
      set_error("mass: bad storage");
      return D2Q9_E_ARG;
  }
  hipLaunchKernelGGL((k_finish<1>), dim3(1), dim3(RED_THREADS), 0, s, work, nb, out);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error(hipGetErrorString(e));
    return (int)e;
  }
  return 0;
}

}  // namespace d2q9
