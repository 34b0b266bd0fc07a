// Momentum exchange on obstacle and wall links (reference:
// pylabolt/parallel/cpu/force_torque_kernels.py:11-141).
//
// One thread per force node.  The post-collision values f*[k] of the node
// come from the current population buffer; the streamed values f_new are
// either read (when the caller has materialised them) or recomputed for
// the node alone with the sweep's own pull + boundary code, so a force
// evaluation costs a pass over the boundary nodes, not over the lattice.
// Per-block FP64 partial sums are added per group in block order.
#include <hip/hip_runtime.h>

#include "d2q9_sweep.h"

using namespace d2q9;

namespace {

constexpr int FORCE_BLOCK = 256;

template <int S>
__global__ __launch_bounds__(FORCE_BLOCK) void k_force_partial(
    const d2q9_lattice L, const typename Codec<S>::raw* __restrict__ fpop,
    const typename Codec<S>::raw* __restrict__ fnew,
    const typename Codec<S>::T* __restrict__ mom, const d2q9_force_node* __restrict__ nodes,
    const d2q9_force_group* __restrict__ groups, double* __restrict__ work) {
  using C = Codec<S>;
  using T = typename C::T;
  __shared__ double red[3][FORCE_BLOCK];
  const int t = (int)threadIdx.x;
  const d2q9_force_node nd = nodes[(int64_t)blockIdx.x * FORCE_BLOCK + t];
  double fx = 0.0, fy = 0.0, tq = 0.0;
  if (nd.mask != 0u) {
    const Consts<T> K = load_consts<T>(L);
    const int64_t c = nd.cell;
    const int i = (int)((c - L.origin) / L.pitch), j = (int)((c - L.origin) % L.pitch);
    T g[9];
    if (fnew == nullptr) {
      gather_cell<S, PHASE_STREAM>(L, K, fpop, mom, i, j, c, L.flags[c], g);
    } else {
#pragma unroll
      for (int k = 0; k < 9; ++k) g[k] = ld<S>(L, K, fnew, i, j, k);
    }
    const d2q9_force_group gr = groups[nd.group];
    double rx = (double)(i - 1) - gr.ref[0], ry = (double)(j - 1) - gr.ref[1];
    if (L.x_periodic) {
      const double nx = (double)L.nx;
      double m = rx;
      if (fabs(rx + nx) < fabs(m)) m = rx + nx;
      if (fabs(rx - nx) < fabs(m)) m = rx - nx;
      rx = m;
    }
    if (L.y_periodic) {
      const double ny = (double)L.ny;
      double m = ry;
      if (fabs(ry + ny) < fabs(m)) m = ry + ny;
      if (fabs(ry - ny) < fabs(m)) m = ry - ny;
      ry = m;
    }
    for (int k = 1; k < 9; ++k) {
      if (!((nd.mask >> k) & 1u)) continue;
      const int kb = INV[k];
      double a = (double)ld<S>(L, K, fpop, i, j, k);
      double b = (double)g[kb];
      if constexpr (C::dev) {
        a += L.w[k];
        b += L.w[kb];
      }
      const double vx = a * (double)CX[k] - b * (double)CX[kb];
      const double vy = a * (double)CY[k] - b * (double)CY[kb];
      fx += vx;
      fy += vy;
      if (gr.torque) tq += rx * vy - ry * vx;
    }
  }
  red[0][t] = fx;
  red[1][t] = fy;
  red[2][t] = tq;
  __syncthreads();
  for (int h = FORCE_BLOCK / 2; h > 0; h >>= 1) {
    if (t < h) {
      red[0][t] += red[0][t + h];
      red[1][t] += red[1][t + h];
      red[2][t] += red[2][t + h];
    }
    __syncthreads();
  }
  if (t < 3) work[(int64_t)blockIdx.x * 3 + t] = red[t][0];
}

__global__ void k_force_finish(const d2q9_force_group* __restrict__ groups, int n_groups,
                               const double* __restrict__ work, double* __restrict__ out) {
  const int g = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (g >= n_groups) return;
  const d2q9_force_group gr = groups[g];
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  for (int b = gr.first_block; b < gr.first_block + gr.n_blocks; ++b) {
    s0 += work[(int64_t)b * 3];
    s1 += work[(int64_t)b * 3 + 1];
    s2 += work[(int64_t)b * 3 + 2];
  }
  out[3 * g] = s0;
  out[3 * g + 1] = s1;
  out[3 * g + 2] = s2;
}

}  // namespace

extern "C" int d2q9_momentum_exchange(const d2q9_lattice* L, const void* f_pop,
                                      const void* f_pop_new, const void* mom,
                                      const d2q9_force_node* nodes, int64_t n_nodes,
                                      const d2q9_force_group* groups, int n_groups,
                                      double* work, double* out, void* stream) {
  if (!L || !f_pop || !mom || !out || (n_nodes > 0 && (!nodes || !groups || !work))) {
    set_error("momentum exchange: null buffer");
    return D2Q9_E_ARG;
  }
  if (n_nodes < 0 || n_nodes % FORCE_BLOCK != 0 || n_groups < 0 ||
      n_nodes / FORCE_BLOCK > 0x7fffffff) {
    set_error("momentum exchange: n_nodes must be a multiple of 256");
    return D2Q9_E_ARG;
  }
  if (L->storage < D2Q9_STORE_F64 || L->storage > D2Q9_STORE_BF16DEV) {
    set_error("momentum exchange: bad storage");
    return D2Q9_E_ARG;
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (n_groups == 0) return 0;
  const unsigned nb = (unsigned)(n_nodes / FORCE_BLOCK);
  if (nb > 0) {
    switch (L->storage) {
#define D2Q9_FORCE(SV)                                                                    \
  case SV:                                                                                \
    hipLaunchKernelGGL((k_force_partial<SV>), dim3(nb), dim3(FORCE_BLOCK), 0, s, *L,      \
                       static_cast<const Codec<SV>::raw*>(f_pop),                         \
                       static_cast<const Codec<SV>::raw*>(f_pop_new),                     \
                       static_cast<const Codec<SV>::T*>(mom), nodes, groups, work);       \
    break;
      D2Q9_FORCE(D2Q9_STORE_F64)
      D2Q9_FORCE(D2Q9_STORE_F32)
      D2Q9_FORCE(D2Q9_STORE_F16DEV)
      D2Q9_FORCE(D2Q9_STORE_BF16DEV)
#undef D2Q9_FORCE
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
      set_error(hipGetErrorString(e));
      return (int)e;
    }
  }
  hipLaunchKernelGGL(k_force_finish, dim3((n_groups + 63) / 64), dim3(64), 0, s, groups,
                     n_groups, work, out);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error(hipGetErrorString(e));
    return (int)e;
  }
  return 0;
}
