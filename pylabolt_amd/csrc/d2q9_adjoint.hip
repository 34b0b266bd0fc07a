// Reverse-mode (adjoint) step of the FP64 D2Q9 BGK time step with Guo
// forcing and a fixed per-cell force field.
//
// Forward step at a fluid cell y (post-collision state p -> p'):
//   g = S(p)[y]                       pull streaming + bounce-back rules
//   rho = sum g,  u = (sum c g + F/2) / rho
//   p'_k = (1 - om) g_k + om feq_k(rho, u) + (1 - om/2) Fbar_k(u, F)
// The collision depends on g only through (rho, u), so with
//   lam' = dL/dp'  and  G_u = dL/du of this step (the loss seed),
//   A_rho = om sum_k lam'_k dfeq_k/drho
//   A_u   = sum_k lam'_k (om dfeq_k/du + (1 - om/2) dFbar_k/du) + G_u
//   mu_k  = dL/dg_k = (1 - om) lam'_k + A_rho + inv (c_k - u) . A_u
//   dL/dF += (1 - om/2) sum_k lam'_k dFbar_k/dF + (inv / 2) A_u
// with inv = 1 / (rho + float_min); then lam = S^T mu.  Streaming with
// bounce-back walls, periodic wraps and static solids is a permutation of
// the fluid populations, so S^T is a scatter in which every target is
// written exactly once: mu_d of cell y goes to the population it was pulled
// from -- (y - c_d, d) wrapped, or (y, inv(d)) where a bounce-back rule or a
// solid neighbour supplied it.  The caller guarantees that precondition
// (no moving walls, pressure or zero-gradient rules, no moving solids).
//
// The constants (inv_cs_2, inv_cs_4, weights) are the forward pass's, so the
// gradient is the exact derivative of the discrete FP64 map.
#include <hip/hip_runtime.h>

#include <string>

#include "d2q9_device.h"

using namespace d2q9;

namespace {

int afail(int code, const char* what) {
  set_error(what);
  return code;
}

constexpr int B64 = 64;

template <int FORCING, bool SCATTER>
__global__ __launch_bounds__(256) void k_adjoint(
    const d2q9_lattice L, const double* __restrict__ lam_src,
    double* __restrict__ lam_dst, const double* __restrict__ mom,
    const double* __restrict__ seed_u, double* __restrict__ grad_f) {
  const int64_t n = (int64_t)L.nx * L.ny;
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const int i = 1 + (int)(e / L.ny), j = 1 + (int)(e % L.ny);
  const int64_t c = cell_index(L, i, j);
  const uint8_t fl = L.flags[c];
  if (!(fl & D2Q9_CELL_FLUID)) return;
  const int64_t P = L.plane;
  const double om = L.omega, a = L.inv_cs_2, i4 = L.inv_cs_4;
  const double b = 0.5 * i4, c0 = 0.5 * a;
  const double* F = static_cast<const double*>(L.force_field);
  const double fx = F[c], fy = F[P + c];
  const double rho = mom[c], ux = mom[P + c], uy = mom[2 * P + c];
  const double inv = 1.0 / (rho + L.float_min);
  const double u2 = ux * ux + uy * uy;
  double lam[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) lam[k] = lam_src[pop_index<B64>(L, i, j, k)];
  double a_rho = 0.0, aux = 0.0, auy = 0.0, gfx = 0.0, gfy = 0.0;
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    const double cx = CX[k], cy = CY[k], w = L.w[k];
    const double cu = cx * ux + cy * uy;
    // d feq / d rho at fixed u, d feq / d u
    const double dfr = w * (1.0 + a * cu + b * cu * cu - c0 * u2);
    const double dfu_x = w * rho * (a * cx + 2.0 * b * cu * cx - 2.0 * c0 * ux);
    const double dfu_y = w * rho * (a * cy + 2.0 * b * cu * cy - 2.0 * c0 * uy);
    a_rho += lam[k] * om * dfr;
    aux += lam[k] * om * dfu_x;
    auy += lam[k] * om * dfu_y;
    if constexpr (FORCING == D2Q9_GUO_LINEAR) {
      const double h = (1.0 - 0.5 * om) * lam[k] * w * a;
      gfx += h * cx;
      gfy += h * cy;
    } else if constexpr (FORCING == D2Q9_GUO_SECOND_ORDER) {
      const double h = (1.0 - 0.5 * om) * lam[k] * w;
      const double cf = cx * fx + cy * fy;
      // dFbar/du = w (-a F + i4 (c.F) c)
      aux += h * (-a * fx + i4 * cf * cx);
      auy += h * (-a * fy + i4 * cf * cy);
      // dFbar/dF = w (a (c - u) + i4 (c.u) c)
      gfx += h * (a * (cx - ux) + i4 * cu * cx);
      gfy += h * (a * (cy - uy) + i4 * cu * cy);
    }
  }
  if (seed_u) {
    aux += seed_u[c];
    auy += seed_u[P + c];
  }
  grad_f[c] += gfx + 0.5 * inv * aux;
  grad_f[P + c] += gfy + 0.5 * inv * auy;
  double mu[9];
#pragma unroll
  for (int k = 0; k < 9; ++k)
    mu[k] = (1.0 - om) * lam[k] + a_rho + inv * ((CX[k] - ux) * aux + (CY[k] - uy) * auy);
  if constexpr (!SCATTER) {
#pragma unroll
    for (int k = 0; k < 9; ++k) lam_dst[pop_index<B64>(L, i, j, k)] = mu[k];
    return;
  } else {
    lam_dst[pop_index<B64>(L, i, j, 0)] = mu[0];
    uint16_t rules[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (fl & D2Q9_CELL_BC) {
      const d2q9_bc_node node = L.bc_nodes[L.bc_slot[c]];
#pragma unroll
      for (int d = 0; d < 9; ++d) rules[d] = node.rule[d];
    }
#pragma unroll
    for (int d = 1; d < 9; ++d) {
      int xn = i - CX[d], yn = j - CY[d];
      bool bounce = (rules[d] & 15) == D2Q9_RULE_BB;
      if (!bounce && (fl & D2Q9_CELL_NEAR_SOLID))
        bounce = (L.flags[cell_index(L, xn, yn)] & D2Q9_CELL_SOLID) != 0;
      if (bounce) {
        lam_dst[pop_index<B64>(L, i, j, INV[d])] = mu[d];
        continue;
      }
      if (L.x_periodic) {
        if (xn == 0) xn = L.nx;
        else if (xn == L.nx + 1) xn = 1;
      }
      if (L.y_periodic) {
        if (yn == 0) yn = L.ny;
        else if (yn == L.ny + 1) yn = 1;
      }
      if (xn < 1 || xn > L.nx || yn < 1 || yn > L.ny) continue;  // from a ghost: no dependence
      lam_dst[pop_index<B64>(L, xn, yn, d)] = mu[d];
    }
  }
}

// dL/du0 (and dL/drho0) from mu_1 = dL/dg_0, g_0 = feq(rho0, u0).
__global__ __launch_bounds__(256) void k_adjoint_init(const d2q9_lattice L,
                                                      const double* __restrict__ mu,
                                                      const double* __restrict__ mom0,
                                                      double* __restrict__ grad_init) {
  const int64_t n = (int64_t)L.nx * L.ny;
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const int i = 1 + (int)(e / L.ny), j = 1 + (int)(e % L.ny);
  const int64_t c = cell_index(L, i, j);
  const int64_t P = L.plane;
  if (!(L.flags[c] & D2Q9_CELL_FLUID)) {
    grad_init[c] = 0.0;
    grad_init[P + c] = 0.0;
    grad_init[2 * P + c] = 0.0;
    return;
  }
  const double a = L.inv_cs_2, b = 0.5 * L.inv_cs_4, c0 = 0.5 * L.inv_cs_2;
  const double rho = mom0[c], ux = mom0[P + c], uy = mom0[2 * P + c];
  const double u2 = ux * ux + uy * uy;
  double gr = 0.0, gx = 0.0, gy = 0.0;
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    const double m = mu[pop_index<B64>(L, i, j, k)];
    const double cx = CX[k], cy = CY[k], w = L.w[k];
    const double cu = cx * ux + cy * uy;
    gr += m * w * (1.0 + a * cu + b * cu * cu - c0 * u2);
    gx += m * w * rho * (a * cx + 2.0 * b * cu * cx - 2.0 * c0 * ux);
    gy += m * w * rho * (a * cy + 2.0 * b * cu * cy - 2.0 * c0 * uy);
  }
  grad_init[c] = gr;
  grad_init[P + c] = gx;
  grad_init[2 * P + c] = gy;
}

int check(const d2q9_lattice* L) {
  if (!L) return afail(D2Q9_E_ARG, "null lattice descriptor");
  if (L->storage != D2Q9_STORE_F64 || L->cpl != 1)
    return afail(D2Q9_E_UNSUPPORTED, "the adjoint runs on FP64 storage");
  if (L->collision != D2Q9_BGK) return afail(D2Q9_E_UNSUPPORTED, "the adjoint needs BGK");
  if (L->force_source != D2Q9_FSRC_FIELD || !L->force_field)
    return afail(D2Q9_E_UNSUPPORTED, "the adjoint needs a per-cell force field");
  if (!L->flags) return afail(D2Q9_E_ARG, "null flags");
  return 0;
}

}  // namespace

extern "C" {

int d2q9_adjoint_step(const d2q9_lattice* L, int streamed, const void* lam_src,
                      void* lam_dst, const void* mom, const void* seed_u,
                      void* grad_force, void* stream) {
  if (int r = check(L)) return r;
  if (!lam_src || !lam_dst || !mom || !grad_force || lam_src == lam_dst)
    return afail(D2Q9_E_ARG, "adjoint buffers: null or aliased");
  const int64_t n = (int64_t)L->nx * L->ny;
  const unsigned blocks = (unsigned)((n + 255) / 256);
  hipStream_t s = static_cast<hipStream_t>(stream);
  auto ls = static_cast<const double*>(lam_src);
  auto ld = static_cast<double*>(lam_dst);
  auto m = static_cast<const double*>(mom);
  auto su = static_cast<const double*>(seed_u);
  auto gf = static_cast<double*>(grad_force);
#define D2Q9_ADJ(FO)                                                                  \
  if (streamed)                                                                       \
    hipLaunchKernelGGL((k_adjoint<FO, true>), dim3(blocks), dim3(256), 0, s, *L, ls, ld, \
                       m, su, gf);                                                    \
  else                                                                                \
    hipLaunchKernelGGL((k_adjoint<FO, false>), dim3(blocks), dim3(256), 0, s, *L, ls,  \
                       ld, m, su, gf);
  switch (L->forcing) {
    case D2Q9_FORCING_NONE: D2Q9_ADJ(D2Q9_FORCING_NONE) break;
    case D2Q9_GUO_LINEAR: D2Q9_ADJ(D2Q9_GUO_LINEAR) break;
    case D2Q9_GUO_SECOND_ORDER: D2Q9_ADJ(D2Q9_GUO_SECOND_ORDER) break;
    default: return afail(D2Q9_E_ARG, "bad forcing");
  }
#undef D2Q9_ADJ
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error(hipGetErrorString(e));
    return (int)e;
  }
  return 0;
}

int d2q9_adjoint_init_grad(const d2q9_lattice* L, const void* mu, const void* mom0,
                           void* grad_init, void* stream) {
  if (int r = check(L)) return r;
  if (!mu || !mom0 || !grad_init) return afail(D2Q9_E_ARG, "null buffer");
  const int64_t n = (int64_t)L->nx * L->ny;
  hipLaunchKernelGGL(k_adjoint_init, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     static_cast<hipStream_t>(stream), *L, static_cast<const double*>(mu),
                     static_cast<const double*>(mom0), static_cast<double*>(grad_init));
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error(hipGetErrorString(e));
    return (int)e;
  }
  return 0;
}

}  // extern "C"
