// Device building blocks of the D2Q9 engine for gfx950 (CDNA4).
//
// Shared by the single-lattice sweep (d2q9_step.hip), the batched persistent
// sweep (d2q9_batched.hip) and the adjoint (d2q9_adjoint.hip).
//
// FP64 arithmetic follows the reference kernels operation by operation, in
// the association Python gives them (left to right), so that the FP64 path
// reproduces the reference bit for bit when compiled with
// -ffp-contract=off:
//   equilibrium  pylabolt/parallel/cpu/equilibrium_kernels.py:28-35
//   moments      pylabolt/parallel/cpu/compute_fields_kernels.py:22-65
//   BGK + Guo    pylabolt/parallel/cpu/collision_kernels.py:37-153
//   pull stream  pylabolt/parallel/cpu/streaming_kernels.py:29-47
//   boundaries   pylabolt/parallel/cpu/fluid_boundary_kernels.py:5-152
// The MRT operator, zero-gradient rule and per-cell force source have no
// reference kernel; their FP64 operation order is the one
// oracle/d2q9_oracle.py defines.
//
// The reduced-precision forms (FP32 compute, FP32/FP16/BF16 storage) hold
// and store deviations d_k = f_k - w_k and compute moments, equilibrium and
// collision in deviation form, which keeps the small velocity-carrying part
// of each population out of the rounding of the O(w_k) rest value.
//
// Population layout (AoSoA): padded cell (i, j), direction k lives at
//   ((i * nbj + jj / B) * 9 + k) * B + jj % B,   jj = j + B - 1,
// B = 64 * cpl cells per block (one wave covers one block of one row: each
// lane owns cpl consecutive cells), so a wave's loads and stores of one
// direction are one contiguous 512 B (FP64) segment and the nine
// directions of a block are adjacent.  Per-cell side arrays (flags, boundary
// slots, moments, force field) use the pitched layout
//   origin + i * pitch + j.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <hip/hip_bf16.h>
#include <stdint.h>

#include <type_traits>

#include "../../include/d2q9.h"

namespace d2q9 {

// Text returned by d2q9_last_error() (per host thread); defined in
// d2q9_step.hip.
void set_error(const char* what);

// Velocity set, reference ordering (base/lattice.py:50-59).
__device__ constexpr int CX[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
__device__ constexpr int CY[9] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
__device__ constexpr int INV[9] = {0, 3, 4, 1, 2, 7, 8, 5, 6};

// Reference MRT moment matrix (base/collision_operator.py:147-157) and the
// squared row norms that give M^-1 = M^T diag(1/norm).
__device__ constexpr int MM[9][9] = {
    {1, 1, 1, 1, 1, 1, 1, 1, 1},   {-4, -1, -1, -1, -1, 2, 2, 2, 2},
    {4, -2, -2, -2, -2, 1, 1, 1, 1}, {0, 1, 0, -1, 0, 1, -1, -1, 1},
    {0, -2, 0, 2, 0, 1, -1, -1, 1}, {0, 0, 1, 0, -1, 1, 1, -1, -1},
    {0, 0, -2, 0, 2, 1, 1, -1, -1}, {0, 1, -1, 1, -1, 0, 0, 0, 0},
    {0, 0, 0, 0, 0, 1, -1, 1, -1}};
__device__ constexpr int ROWNORM[9] = {9, 36, 36, 6, 12, 6, 12, 4, 4};

// ---------------------------------------------------------------------------
// Scalar constants in the compute type, loaded once per thread from the
// descriptor (double fields for FP64, the float mirror otherwise).
// ---------------------------------------------------------------------------
template <typename T> struct Consts {
  T w[9], rates[9], scale[9], iscale[9];
  T om, ics2, ics4, fmin, gx, gy;
  T nzero;  // -0, computed at run time (see Codec<D2Q9_STORE_F16DEV>::dec)
};

template <typename T>
__device__ __forceinline__ Consts<T> load_consts(const d2q9_lattice& L);

template <>
__device__ __forceinline__ Consts<double> load_consts<double>(const d2q9_lattice& L) {
  Consts<double> K;
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    K.w[k] = L.w[k];
    K.rates[k] = L.rates[k];
    K.scale[k] = 1.0;
    K.iscale[k] = 1.0;
  }
  K.om = L.omega;
  K.ics2 = L.inv_cs_2;
  K.ics4 = L.inv_cs_4;
  K.fmin = L.float_min;
  K.gx = L.gravity[0];
  K.gy = L.gravity[1];
  K.nzero = -0.0;
  return K;
}

template <>
__device__ __forceinline__ Consts<float> load_consts<float>(const d2q9_lattice& L) {
  Consts<float> K;
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    K.w[k] = L.cf.w[k];
    K.rates[k] = L.cf.rates[k];
    K.scale[k] = L.cf.scale[k];
    K.iscale[k] = L.cf.inv_scale[k];
  }
  K.om = L.cf.omega;
  K.ics2 = L.cf.inv_cs_2;
  K.ics4 = L.cf.inv_cs_4;
  K.fmin = L.cf.float_min;
  K.gx = L.cf.gravity[0];
  K.gy = L.cf.gravity[1];
  // -0 * (a positive scale) = -0; formed at run time, so the compiler keeps
  // the additions of it (x + -0 = x exactly)
  K.nzero = -0.0f * K.iscale[0];
  return K;
}

// ---------------------------------------------------------------------------
// Storage codecs: raw HBM element <-> compute value.
//   dev == false: the value is f_k (FP64 reference path)
//   dev == true:  the value is d_k = f_k - w_k; raw = d_k * scale_k
// CPL: cells per lane (FP64 1: 8 bytes per lane and direction; FP32 2:
// 8 bytes; 16-bit 2: 4 bytes).
// ---------------------------------------------------------------------------
// 32-bit integer hash ("lowbias32" finaliser, C. Wellons, 2018): xor-shift
// and multiply rounds with full avalanche.
__device__ __forceinline__ uint32_t mix32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}

// Random bits for rounding value `b` (float bits) of direction k of cell
// `salt`: a function of the data only, so a run is reproducible bit for
// bit, while the rounding of a slowly changing field is decorrelated from
// step to step.  The cell's key is hashed once per cell (cell_key); per
// value one 24-bit multiply (full rate) spreads the low 24 bits of
// value ^ key into the upper bits of the product, which the conversions
// below use: bits 16..31 for bfloat16, bits 19..31 for half.
__device__ __forceinline__ uint32_t cell_key(uint32_t salt) { return mix32(salt); }
__device__ __forceinline__ uint32_t round_bits(uint32_t b, int k, uint32_t key) {
  const uint32_t x = b ^ (key + (uint32_t)k * 0x632be5abu);
  return __umul24(x, 0x9e3779u);
}

// gfx950's stochastic-rounding conversions.  v_cvt_sr_bf16_f32 adds bits
// 16..31 of the random operand below the 16 kept bits of the float and
// truncates; v_cvt_sr_f16_f32 adds bits 19..31 below half's 10 kept
// mantissa bits over half's normal range and rounds stochastically on the
// subnormal grid below it.  Each writes one half of a 32-bit register
// (word_sel), so a pair of values packs with no extra instruction.
typedef _Float16 half2_sr __attribute__((ext_vector_type(2)));
typedef __bf16 bf162_sr __attribute__((ext_vector_type(2)));

__device__ __forceinline__ uint32_t sr_half2(float a0, uint32_t r0, float a1, uint32_t r1) {
  half2_sr z = {};
  z = __builtin_amdgcn_cvt_sr_f16_f32(z, a0, r0, false);
  z = __builtin_amdgcn_cvt_sr_f16_f32(z, a1, r1, true);
  uint32_t w;
  __builtin_memcpy(&w, &z, 4);
  return w;
}
__device__ __forceinline__ uint32_t sr_bf162(float a0, uint32_t r0, float a1, uint32_t r1) {
  bf162_sr z = {};
  z = __builtin_amdgcn_cvt_sr_bf16_f32(z, a0, r0, false);
  z = __builtin_amdgcn_cvt_sr_bf16_f32(z, a1, r1, true);
  uint32_t w;
  __builtin_memcpy(&w, &z, 4);
  return w;
}

template <int S> struct Codec;

template <> struct Codec<D2Q9_STORE_F64> {
  using raw = double;
  using T = double;
  static constexpr bool dev = false;
  static constexpr int CPL = 1;
  static __device__ __forceinline__ T dec(raw r, int, const Consts<T>&) { return r; }
  static __device__ __forceinline__ raw enc(T v, int, const Consts<T>&) { return v; }
  static __device__ __forceinline__ raw encs(T v, int, uint32_t, const Consts<T>&) { return v; }
  static __device__ __forceinline__ double f64(raw r, int, const d2q9_lattice&) { return r; }
};

// FP32 and BF16 store d unscaled (dev_scale must be 1); FP16 scales it.
template <> struct Codec<D2Q9_STORE_F32> {
  using raw = float;
  using T = float;
  static constexpr bool dev = true;
  static constexpr int CPL = 2;
  static __device__ __forceinline__ T dec(raw r, int, const Consts<T>&) { return r; }
  static __device__ __forceinline__ raw enc(T v, int, const Consts<T>&) { return v; }
  static __device__ __forceinline__ raw encs(T v, int, uint32_t, const Consts<T>&) { return v; }
  static __device__ __forceinline__ double f64(raw r, int k, const d2q9_lattice& L) {
    return L.w[k] + (double)r;
  }
};

// cells per lane of the 16-bit forms (one lane moves 4 bytes per direction)
constexpr int CPL16 = 2;

template <> struct Codec<D2Q9_STORE_F16DEV> {
  using raw = __half;
  using T = float;
  static constexpr bool dev = true;
  static constexpr int CPL = CPL16;
  // h * iscale + (-0) is h * iscale exactly; written as a fused multiply-
  // add with a run-time addend it becomes one v_fma_mix_f32, which reads
  // the half operand (either half of a register) directly.
  static __device__ __forceinline__ T dec(raw r, int k, const Consts<T>& K) {
    return __builtin_fmaf(__half2float(r), K.iscale[k], K.nzero);
  }
  static __device__ __forceinline__ raw enc(T v, int k, const Consts<T>& K) {
    return __float2half_rn(v * K.scale[k]);
  }
  // Stochastic rounding to half of the scaled deviation.
  static __device__ __forceinline__ raw encs(T v, int k, uint32_t key, const Consts<T>& K) {
    const float a = v * K.scale[k];
    half2_sr z = {};
    z = __builtin_amdgcn_cvt_sr_f16_f32(z, a, round_bits(__float_as_uint(a), k, key), false);
    _Float16 h = z[0];
    raw r;
    __builtin_memcpy(&r, &h, 2);
    return r;
  }
  // Two cells' values of direction k, packed (first cell in the low half).
  static __device__ __forceinline__ uint32_t encs2(T v0, T v1, int k, uint32_t key0,
                                                   uint32_t key1, const Consts<T>& K) {
    const float a0 = v0 * K.scale[k], a1 = v1 * K.scale[k];
    return sr_half2(a0, round_bits(__float_as_uint(a0), k, key0), a1,
                    round_bits(__float_as_uint(a1), k, key1));
  }
  static __device__ __forceinline__ double f64(raw r, int k, const d2q9_lattice& L) {
    return L.w[k] + (double)__half2float(r) / L.dev_scale[k];
  }
};

template <> struct Codec<D2Q9_STORE_BF16DEV> {
  using raw = __hip_bfloat16;
  using T = float;
  static constexpr bool dev = true;
  static constexpr int CPL = CPL16;
  static __device__ __forceinline__ T dec(raw r, int, const Consts<T>&) {
    return __bfloat162float(r);
  }
  static __device__ __forceinline__ raw enc(T v, int, const Consts<T>&) {
    return __float2bfloat16(v);
  }
  // Stochastic rounding to bfloat16 (the upper 16 bits of a float).
  static __device__ __forceinline__ raw encs(T v, int k, uint32_t key, const Consts<T>&) {
    bf162_sr z = {};
    z = __builtin_amdgcn_cvt_sr_bf16_f32(z, v, round_bits(__float_as_uint(v), k, key), false);
    __bf16 h = z[0];
    raw r;
    __builtin_memcpy(&r, &h, 2);
    return r;
  }
  static __device__ __forceinline__ uint32_t encs2(T v0, T v1, int k, uint32_t key0,
                                                   uint32_t key1, const Consts<T>&) {
    return sr_bf162(v0, round_bits(__float_as_uint(v0), k, key0), v1,
                    round_bits(__float_as_uint(v1), k, key1));
  }
  static __device__ __forceinline__ double f64(raw r, int k, const d2q9_lattice& L) {
    return L.w[k] + (double)__bfloat162float(r);
  }
};

// Encoded value of a zero population (f = 0): ghost and solid cells.
template <int S>
__device__ __forceinline__ typename Codec<S>::raw enc_zero(int k, const Consts<typename Codec<S>::T>& K) {
  using T = typename Codec<S>::T;
  if constexpr (Codec<S>::dev) return Codec<S>::enc(-K.w[k], k, K);
  else return Codec<S>::enc((T)0, k, K);
}

// ---------------------------------------------------------------------------
// Layout.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int64_t cell_index(const d2q9_lattice& L, int i, int j) {
  return (int64_t)L.origin + (int64_t)i * L.pitch + j;
}

template <int B>
__device__ __forceinline__ int64_t pop_index(const d2q9_lattice& L, int i, int j, int k) {
  const int jj = j + B - 1;
  return (((int64_t)i * L.nbj + jj / B) * 9 + k) * B + (jj % B);
}

// ---------------------------------------------------------------------------
// Equilibrium and Guo source.
// ---------------------------------------------------------------------------
// Reference form: feq_k = (w rho) (((1 + ics2 cu) + ((0.5 ics4) cu) cu) - (0.5 ics2) u2)
// c_k . u with the zero components of c_k left out and the unit ones as a
// sign: for finite u the same bits as cx*ux + cy*uy (a product by +-1 is
// exact, and a +-0 term leaves a finite sum unchanged but for the sign of
// a zero, which only ever enters through 1 + ics2 cu and cu*cu), without
// the products by zero IEEE arithmetic would keep.
template <typename T>
__device__ __forceinline__ T cdot(int k, T ux, T uy) {
  if (CX[k] == 0 && CY[k] == 0) return (T)0;
  if (CX[k] == 0) return CY[k] > 0 ? uy : -uy;
  const T a = CX[k] > 0 ? ux : -ux;
  if (CY[k] == 0) return a;
  return CY[k] > 0 ? a + uy : a - uy;
}

template <typename T>
__device__ __forceinline__ T feq_k(int k, T rho, T ux, T uy, T u2, T wk,
                                   T ics2, T ics4) {
  T cu = cdot<T>(k, ux, uy);
  return (wk * rho) *
         (((T)1 + ics2 * cu + ((T)0.5 * ics4) * cu * cu) - ((T)0.5 * ics2) * u2);
}

// Deviation form: feq_k - w_k = w (drho + rho (ics2 cu + 0.5 ics4 cu^2 - 0.5 ics2 u2))
template <typename T>
__device__ __forceinline__ T deq_k(int k, T drho, T rho, T ux, T uy, T u2, T wk,
                                   T ics2, T ics4) {
  T cu = cdot<T>(k, ux, uy);
  return wk * (drho + rho * (ics2 * cu + ((T)0.5 * ics4) * cu * cu - ((T)0.5 * ics2) * u2));
}

// Guo source without the (1 - omega/2) factor.
template <typename T, int FORCING>
__device__ __forceinline__ T guo_k(int k, T ux, T uy, T fx, T fy, T wk,
                                   T ics2, T ics4) {
  // zero components left out as in cdot: c - u is -u for c = 0, and the
  // (cu c) ics4 term adds a zero
  if constexpr (FORCING == D2Q9_GUO_LINEAR) {
    return (wk * cdot<T>(k, fx, fy)) * ics2;
  } else {
    T cu = cdot<T>(k, ux, uy);
    T const_x = ((T)CX[k] - ux) * ics2;
    if (CX[k] != 0) const_x = const_x + (CX[k] > 0 ? cu : -cu) * ics4;
    T const_y = ((T)CY[k] - uy) * ics2;
    if (CY[k] != 0) const_y = const_y + (CY[k] > 0 ? cu : -cu) * ics4;
    return wk * (const_x * fx + const_y * fy);
  }
}

// Moments of the pre-collision values g (f, or d in deviation form):
// rho = sum g (from 0) [+ 1 in deviation form], F = rho * gravity or the
// per-cell field, u = (sum c g) * inv + (0.5 F) * inv, inv = 1/(rho + float_min).
// Float moments (deviation form): opposite-direction sums, written-out
// fused multiply-adds and the hardware reciprocal.
template <int FORCING>
__device__ __forceinline__ void moments_f32(const float (&g)[9], const d2q9_lattice& L,
                                            const Consts<float>& K, float ffx, float ffy,
                                            float& drho, float& rho, float& ux, float& uy,
                                            float& fx, float& fy) {
  // pairs of opposite directions first: fewer dependent adds
  const float s13 = g[1] + g[3], s24 = g[2] + g[4], s57 = g[5] + g[7], s68 = g[6] + g[8];
  const float d13 = g[1] - g[3], d24 = g[2] - g[4], d57 = g[5] - g[7], d68 = g[6] - g[8];
  drho = (g[0] + s13) + (s24 + (s57 + s68));
  rho = 1.0f + drho;
  const bool field = L.force_source == D2Q9_FSRC_FIELD;
  if (field) {
    fx = ffx;
    fy = ffy;
  } else {
    fx = rho * K.gx;
    fy = rho * K.gy;
  }
  // c_x: +1 for 1,5,8; -1 for 3,6,7.  c_y: +1 for 2,5,6; -1 for 4,7,8.
  const float sx = d13 + (d57 - d68);
  const float sy = d24 + (d57 + d68);
  const float inv = __builtin_amdgcn_rcpf(rho + K.fmin);
  if (FORCING != D2Q9_FORCING_NONE || field) {
    ux = __builtin_fmaf(0.5f, fx, sx) * inv;
    uy = __builtin_fmaf(0.5f, fy, sy) * inv;
  } else {
    ux = sx * inv;
    uy = sy * inv;
  }
}

template <typename T, int FORCING, bool DEV>
__device__ __forceinline__ void moments(const T (&g)[9], const d2q9_lattice& L,
                                        const Consts<T>& K, T ffx, T ffy,
                                        T& drho, T& rho, T& ux, T& uy, T& fx,
                                        T& fy) {
  if constexpr (std::is_same<T, float>::value && DEV) {
    moments_f32<FORCING>(g, L, K, ffx, ffy, drho, rho, ux, uy, fx, fy);
    return;
  }
  T r = (T)0;
#pragma unroll
  for (int k = 0; k < 9; ++k) r += g[k];
  if constexpr (DEV) {
    drho = r;
    rho = (T)1 + r;
  } else {
    drho = (T)0;
    rho = r;
  }
  const bool field = L.force_source == D2Q9_FSRC_FIELD;
  if (field) {
    fx = ffx;
    fy = ffy;
  } else {
    fx = rho * K.gx;
    fy = rho * K.gy;
  }
  // sum c_k g_k from 0 in k order; zero terms left out (a partial sum that
  // starts at +0 is never -0, so adding +-0 would leave it unchanged)
  T sx = (T)0, sy = (T)0;
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    if (CX[k] != 0) sx = CX[k] > 0 ? sx + g[k] : sx - g[k];
    if (CY[k] != 0) sy = CY[k] > 0 ? sy + g[k] : sy - g[k];
  }
  // FP64 divides as the reference does; the float forms use the
  // hardware reciprocal (1 ulp)
  T inv;
  if constexpr (std::is_same<T, double>::value)
    inv = (T)1 / (rho + K.fmin);
  else
    inv = __builtin_amdgcn_rcpf(rho + K.fmin);
  if (FORCING != D2Q9_FORCING_NONE || field) {
    ux = sx * inv + ((T)0.5 * fx) * inv;
    uy = sy * inv + ((T)0.5 * fy) * inv;
  } else {
    // gravity is zero when forcing is None (base/force_operator.py:58-69),
    // so the force term adds exactly zero.
    ux = sx * inv;
    uy = sy * inv;
  }
}

// Float BGK (deviation form) over the four opposite-direction pairs.  With
// cu_kb = -cu_k the relaxed equilibrium om geq splits into a part shared by
// the pair, A = om w (drho + rho (0.5 ics4 cu^2 - 0.5 ics2 u^2)), and one of
// opposite sign, B = om w rho ics2 cu; the Guo source (times 1 - om/2)
// likewise into (1 - om/2) w (ics4 cu (c.F) - ics2 (u.F)) (second order
// only) and (1 - om/2) w ics2 (c.F).  Then g_k <- (1 - om) g_k + (A + B),
// g_kb <- (1 - om) g_kb + (A - B), with om w rho and om w drho formed once
// per cell and weight.
template <int FORCING>
__device__ __forceinline__ void collide_bgk_pairs(float (&g)[9], const Consts<float>& K,
                                                  float drho, float rho, float ux, float uy,
                                                  float fx, float fy) {
  // fused multiply-adds are written out (never left to the compiler), so
  // every kernel build computes the same values
  const float om = K.om;
  const float omc = 1.0f - om;
  const float ics2 = K.ics2;
  const float hics4 = 0.5f * K.ics4;
  const float hu2 = (0.5f * ics2) * __builtin_fmaf(ux, ux, uy * uy);
  const float cf = __builtin_fmaf(-0.5f, om, 1.0f);
  float uf = 0.0f;
  if constexpr (FORCING == D2Q9_GUO_SECOND_ORDER) uf = ics2 * __builtin_fmaf(ux, fx, uy * fy);
  // rest direction (cu = 0, c = 0)
  {
    const float omw = om * K.w[0];
    float a = __builtin_fmaf(-(rho * omw), hu2, drho * omw);
    if constexpr (FORCING == D2Q9_GUO_SECOND_ORDER) a = __builtin_fmaf(cf * K.w[0], -uf, a);
    g[0] = __builtin_fmaf(omc, g[0], a);
  }
  constexpr int PK[4] = {1, 2, 5, 6};
  constexpr int PB[4] = {3, 4, 7, 8};
  // per weight class (axis pairs 0, 1; diagonal pairs 2, 3)
  float rw[2], dw[2], aw[2];
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    const float omw = om * K.w[c ? 5 : 1];
    rw[c] = rho * omw;
    dw[c] = drho * omw;
    aw[c] = rw[c] * ics2;
  }
#pragma unroll
  for (int n = 0; n < 4; ++n) {
    const int k = PK[n], kb = PB[n], c = n >> 1;
    const float cu = (float)CX[k] * ux + (float)CY[k] * uy;
    const float t = __builtin_fmaf(hics4 * cu, cu, -hu2);
    float A = __builtin_fmaf(rw[c], t, dw[c]);
    float B = aw[c] * cu;
    if constexpr (FORCING != D2Q9_FORCING_NONE) {
      const float w = K.w[k];
      const float cF = (float)CX[k] * fx + (float)CY[k] * fy;
      B = __builtin_fmaf((cf * w) * ics2, cF, B);
      if constexpr (FORCING == D2Q9_GUO_SECOND_ORDER)
        A = __builtin_fmaf(cf * w, __builtin_fmaf(K.ics4 * cu, cF, -uf), A);
    }
    g[k] = __builtin_fmaf(omc, g[k], A + B);
    g[kb] = __builtin_fmaf(omc, g[kb], A - B);
  }
}

// BGK collision in place: g <- (1 - om) g + om geq (+ (1 - om/2) Fbar).
template <typename T, int FORCING, bool DEV>
__device__ __forceinline__ void collide_bgk(T (&g)[9], const Consts<T>& K,
                                            T drho, T rho, T ux, T uy, T fx, T fy) {
  if constexpr (std::is_same<T, float>::value && DEV) {
    collide_bgk_pairs<FORCING>(g, K, drho, rho, ux, uy, fx, fy);
    return;
  }
  const T om = K.om;
  const T ics2 = K.ics2, ics4 = K.ics4;
  const T u2 = ux * ux + uy * uy;
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    T wk = K.w[k];
    T eq = DEV ? deq_k<T>(k, drho, rho, ux, uy, u2, wk, ics2, ics4)
               : feq_k<T>(k, rho, ux, uy, u2, wk, ics2, ics4);
    T v = ((T)1 - om) * g[k] + om * eq;
    if constexpr (FORCING != D2Q9_FORCING_NONE) {
      v = v + ((T)1 - (T)0.5 * om) * guo_k<T, FORCING>(k, ux, uy, fx, fy, wk, ics2, ics4);
    }
    g[k] = v;
  }
}

// The nine moments (rows of M) of a population-like vector given as its
// rest value x0 and the sums S and differences D of the opposite pairs
// (1,3), (2,4), (5,7), (6,8).
__device__ __forceinline__ void moments_of_pairs(float x0, const float (&S)[4],
                                                 const float (&D)[4], float (&m)[9]) {
  const float s12 = S[0] + S[1], s34 = S[2] + S[3];
  m[0] = x0 + s12 + s34;
  m[1] = __builtin_fmaf(2.0f, s34, __builtin_fmaf(-4.0f, x0, -s12));
  m[2] = __builtin_fmaf(-2.0f, s12, __builtin_fmaf(4.0f, x0, s34));
  const float a = D[2] - D[3], b = D[2] + D[3];
  m[3] = D[0] + a;
  m[4] = __builtin_fmaf(-2.0f, D[0], a);
  m[5] = D[1] + b;
  m[6] = __builtin_fmaf(-2.0f, D[1], b);
  m[7] = S[0] - S[1];
  m[8] = S[2] - S[3];
}

// Float MRT (deviation form) with the same pair structure: the columns of
// M for k and its opposite kb share the rows 0, 1, 2, 7, 8 and negate the
// rows 3..6, so M (g - geq), M Fbar and the inverse map M^T diag(1/norm)
// all go through pair sums and differences.  Same operator as the FP64
// path, f* = f - M^-1 [S (M f - M feq) - (I - S/2) M Fbar], different
// rounding.
template <int FORCING>
__device__ __forceinline__ void collide_mrt_pairs(float (&g)[9], const Consts<float>& K,
                                                  float drho, float rho, float ux, float uy,
                                                  float fx, float fy) {
  const float ics2 = K.ics2;
  const float hics4 = 0.5f * K.ics4;
  const float hu2 = (0.5f * ics2) * __builtin_fmaf(ux, ux, uy * uy);
  const float rics2 = rho * ics2;
  const float uf = ics2 * __builtin_fmaf(ux, fx, uy * fy);
  constexpr int PK[4] = {1, 2, 5, 6};
  constexpr int PB[4] = {3, 4, 7, 8};
  // non-equilibrium part e = g - geq and the Guo source, as pair sums and
  // differences
  const float e0 = g[0] - K.w[0] * __builtin_fmaf(-rho, hu2, drho);
  float eS[4], eD[4], fS[4], fD[4];
  float f0 = 0.0f;
  if constexpr (FORCING == D2Q9_GUO_SECOND_ORDER) f0 = -(K.w[0] * uf);
#pragma unroll
  for (int n = 0; n < 4; ++n) {
    const int k = PK[n], kb = PB[n];
    const float cu = (float)CX[k] * ux + (float)CY[k] * uy;
    const float w = K.w[k];
    const float sym = __builtin_fmaf(rho, __builtin_fmaf(hics4 * cu, cu, -hu2), drho);
    const float asym = rics2 * cu;
    // geq_k + geq_kb = 2 w sym, geq_k - geq_kb = 2 w asym
    eS[n] = (g[k] + g[kb]) - (2.0f * w) * sym;
    eD[n] = (g[k] - g[kb]) - (2.0f * w) * asym;
    fS[n] = 0.0f;
    fD[n] = 0.0f;
    if constexpr (FORCING != D2Q9_FORCING_NONE) {
      const float cF = (float)CX[k] * fx + (float)CY[k] * fy;
      fD[n] = (2.0f * w) * (ics2 * cF);
      if constexpr (FORCING == D2Q9_GUO_SECOND_ORDER)
        fS[n] = (2.0f * w) * __builtin_fmaf(K.ics4 * cu, cF, -uf);
    }
  }
  float me[9], mf[9];
  moments_of_pairs(e0, eS, eD, me);
  if constexpr (FORCING != D2Q9_FORCING_NONE) moments_of_pairs(f0, fS, fD, mf);
  float gm[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    const float sr = K.rates[i];
    float dm = sr * me[i];
    if constexpr (FORCING != D2Q9_FORCING_NONE)
      dm = __builtin_fmaf(-__builtin_fmaf(-0.5f, sr, 1.0f), mf[i], dm);
    gm[i] = dm * (float)(1.0 / (double)ROWNORM[i]);
  }
  // g -= M^T gm, by columns: rest, then per pair the shared and the
  // opposite-signed part
  g[0] -= __builtin_fmaf(4.0f, gm[2] - gm[1], gm[0]);
  const float base13 = __builtin_fmaf(-2.0f, gm[2], gm[0] - gm[1]);
  const float sym13 = base13 + gm[7], sym24 = base13 - gm[7];
  const float asym13 = __builtin_fmaf(-2.0f, gm[4], gm[3]);
  const float asym24 = __builtin_fmaf(-2.0f, gm[6], gm[5]);
  const float base57 = __builtin_fmaf(2.0f, gm[1], gm[0] + gm[2]);
  const float sym57 = base57 + gm[8], sym68 = base57 - gm[8];
  const float p34 = gm[3] + gm[4], p56 = gm[5] + gm[6];
  const float asym57 = p34 + p56, asym68 = p56 - p34;
  g[1] -= sym13 + asym13;
  g[3] -= sym13 - asym13;
  g[2] -= sym24 + asym24;
  g[4] -= sym24 - asym24;
  g[5] -= sym57 + asym57;
  g[7] -= sym57 - asym57;
  g[6] -= sym68 + asym68;
  g[8] -= sym68 - asym68;
}

// Moments (rows of M) of a population-like vector from its rest value x0
// and the sums S and differences D of the pairs (1,3), (2,4), (5,7),
// (6,8), without fused multiply-adds: the FP64 operation order of oracle
// pair_moments.
template <typename T>
__device__ __forceinline__ void pair_moments(T x0, const T (&S)[4], const T (&D)[4],
                                             T (&m)[9]) {
  const T s12 = S[0] + S[1], s34 = S[2] + S[3];
  const T a = D[2] - D[3], b = D[2] + D[3];
  m[0] = (x0 + s12) + s34;
  m[1] = (((T)-4 * x0) - s12) + (T)2 * s34;
  m[2] = (((T)4 * x0) + s34) - (T)2 * s12;
  m[3] = D[0] + a;
  m[4] = a - (T)2 * D[0];
  m[5] = D[1] + b;
  m[6] = b - (T)2 * D[1];
  m[7] = S[0] - S[1];
  m[8] = S[2] - S[3];
}

// M^T gm by columns (oracle pair_inverse).
template <typename T>
__device__ __forceinline__ void pair_inverse(const T (&gm)[9], T (&out)[9]) {
  out[0] = gm[0] + (T)4 * (gm[2] - gm[1]);
  const T base13 = (gm[0] - gm[1]) - (T)2 * gm[2];
  const T sym13 = base13 + gm[7], sym24 = base13 - gm[7];
  const T asym13 = gm[3] - (T)2 * gm[4], asym24 = gm[5] - (T)2 * gm[6];
  const T base57 = (gm[0] + gm[2]) + (T)2 * gm[1];
  const T sym57 = base57 + gm[8], sym68 = base57 - gm[8];
  const T p34 = gm[3] + gm[4], p56 = gm[5] + gm[6];
  const T asym57 = p34 + p56, asym68 = p56 - p34;
  out[1] = sym13 + asym13;
  out[2] = sym24 + asym24;
  out[3] = sym13 - asym13;
  out[4] = sym24 - asym24;
  out[5] = sym57 + asym57;
  out[6] = sym68 + asym68;
  out[7] = sym57 - asym57;
  out[8] = sym68 - asym68;
}

// MRT with Guo forcing in place, FP64 storage: the operation order of
// oracle mrt_collide,
//   e = g - feq, Fbar = Guo source (pair sums and differences),
//   p = (I - S) M e + (I - S/2) M Fbar,  g = feq + M^-1 p,
// which equals g - M^-1 [S (M g - M feq) - (I - S/2) M Fbar].  The
// populations are consumed pair by pair and feq is evaluated again for the
// result instead of being held across the moments (the compiler is shown
// copies of rho and u it cannot see through, so it cannot merge the two
// evaluations): the live set stays near 24 doubles, which keeps the
// two-step kernel at two workgroups per CU.
template <typename T, int FORCING, bool DEV>
__device__ __forceinline__ void collide_mrt(T (&g)[9], const Consts<T>& K,
                                            T drho, T rho, T ux, T uy, T fx, T fy) {
  static_assert(!DEV, "deviation storage forms use collide_mrt_pairs");
  constexpr bool forced = FORCING != D2Q9_FORCING_NONE;
  constexpr int PK[4] = {1, 2, 5, 6};
  constexpr int PB[4] = {3, 4, 7, 8};
  const T ics2 = K.ics2, ics4 = K.ics4;
  // moments of the Guo source first; the equilibrium part then starts from
  // copies of rho and u that depend on them, so that the two sets of pair
  // sums and differences are never live together
  T mf[9];
  if constexpr (forced) {
    T fS[4], fD[4];
    const T f0 = guo_k<T, FORCING>(0, ux, uy, fx, fy, K.w[0], ics2, ics4);
#pragma unroll
    for (int n = 0; n < 4; ++n) {
      const int k = PK[n], kb = PB[n];
      const T fk = guo_k<T, FORCING>(k, ux, uy, fx, fy, K.w[k], ics2, ics4);
      const T fkb = guo_k<T, FORCING>(kb, ux, uy, fx, fy, K.w[kb], ics2, ics4);
      fS[n] = fk + fkb;
      fD[n] = fk - fkb;
    }
    pair_moments<T>(f0, fS, fD, mf);
    asm volatile("" : "+v"(rho), "+v"(ux), "+v"(uy)
                 : "v"(mf[0]), "v"(mf[1]), "v"(mf[2]), "v"(mf[3]), "v"(mf[4]),
                   "v"(mf[5]), "v"(mf[6]), "v"(mf[7]), "v"(mf[8]));
  }
  const T u2 = ux * ux + uy * uy;
  const T e0 = g[0] - feq_k<T>(0, rho, ux, uy, u2, K.w[0], ics2, ics4);
  T eS[4], eD[4];
#pragma unroll
  for (int n = 0; n < 4; ++n) {
    const int k = PK[n], kb = PB[n];
    const T ek = g[k] - feq_k<T>(k, rho, ux, uy, u2, K.w[k], ics2, ics4);
    const T ekb = g[kb] - feq_k<T>(kb, rho, ux, uy, u2, K.w[kb], ics2, ics4);
    eS[n] = ek + ekb;
    eD[n] = ek - ekb;
  }
  T me[9], gm[9];
  pair_moments<T>(e0, eS, eD, me);
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    const T s = K.rates[i];
    T p = ((T)1 - s) * me[i];
    if constexpr (forced) p = p + ((T)1 - (T)0.5 * s) * mf[i];
    gm[i] = p * (T)(1.0 / (double)ROWNORM[i]);
  }
  T out[9];
  pair_inverse<T>(gm, out);
  T r2 = rho, x2 = ux, y2 = uy;
  asm volatile("" : "+v"(r2), "+v"(x2), "+v"(y2));
  const T u2b = x2 * x2 + y2 * y2;
#pragma unroll
  for (int k = 0; k < 9; ++k)
    g[k] = feq_k<T>(k, r2, x2, y2, u2b, K.w[k], ics2, ics4) + out[k];
  (void)drho;
}

template <typename T, int MODEL, int FORCING, bool DEV>
__device__ __forceinline__ void collide(T (&g)[9], const Consts<T>& K, T drho,
                                        T rho, T ux, T uy, T fx, T fy) {
  if constexpr (MODEL == D2Q9_MRT) {
    if constexpr (std::is_same<T, float>::value && DEV)
      collide_mrt_pairs<FORCING>(g, K, drho, rho, ux, uy, fx, fy);
    else
      collide_mrt<T, FORCING, DEV>(g, K, drho, rho, ux, uy, fx, fy);
  } else
    collide_bgk<T, FORCING, DEV>(g, K, drho, rho, ux, uy, fx, fy);
}

// ---------------------------------------------------------------------------
// Streaming and boundary rules on the general (flagged) path, one cell.
// ---------------------------------------------------------------------------
template <int S>
__device__ __forceinline__ typename Codec<S>::T ld(const d2q9_lattice& L,
                                                   const Consts<typename Codec<S>::T>& K,
                                                   const typename Codec<S>::raw* __restrict__ f,
                                                   int i, int j, int k) {
  constexpr int B = 64 * Codec<S>::CPL;
  return Codec<S>::dec(f[pop_index<B>(L, i, j, k)], k, K);
}

// Value streamed into cell (i, j) in direction k (streaming_kernels.py:35-47
// with the periodic wrap of MPI_kernels.py:34-99 folded into the index):
// from the pull neighbour, or moving-wall bounce-back off a solid neighbour.
template <int S>
__device__ __forceinline__ typename Codec<S>::T pull(
    const d2q9_lattice& L, const Consts<typename Codec<S>::T>& K,
    const typename Codec<S>::raw* __restrict__ f,
    const typename Codec<S>::T* __restrict__ mom, int i, int j, int64_t c,
    uint8_t fl, int k) {
  using T = typename Codec<S>::T;
  int xn = i - CX[k], yn = j - CY[k];
  if (fl & D2Q9_CELL_NEAR_SOLID) {
    int64_t nb = cell_index(L, xn, yn);
    if (L.flags[nb] & D2Q9_CELL_SOLID) {
      const int ki = INV[k];
      T rho_c = mom[c];
      T usx = mom[L.plane + nb], usy = mom[2 * L.plane + nb];
      T temp = ((((T)2 * K.w[ki]) * rho_c) * K.ics2) *
               ((T)CX[ki] * usx + (T)CY[ki] * usy);
      // w[ki] == w[k]: the same expression holds for deviations
      return ld<S>(L, K, f, i, j, ki) - temp;
    }
  }
  if (L.x_periodic) {
    if (xn == 0) xn = L.nx;
    else if (xn == L.nx + 1) xn = 1;
  }
  if (L.y_periodic) {
    if (yn == 0) yn = L.ny;
    else if (yn == L.ny + 1) yn = 1;
  }
  return ld<S>(L, K, f, xn, yn, k);
}

// Overwrite the rule-bearing directions of boundary node (i, j)
// (fluid_boundary_kernels.py:5-152; rules resolved in element order on the
// host, so the last element writing a direction is the one stored).
template <int S>
__device__ __forceinline__ void apply_bc(
    const d2q9_lattice& L, const Consts<typename Codec<S>::T>& K,
    const typename Codec<S>::raw* __restrict__ f,
    const typename Codec<S>::T* __restrict__ mom, int i, int j, int64_t c,
    typename Codec<S>::T (&g)[9]) {
  using T = typename Codec<S>::T;
  const d2q9_bc_node node = L.bc_nodes[L.bc_slot[c]];
  const T ics2 = K.ics2, ics4 = K.ics4;
#pragma unroll
  for (int d = 1; d < 9; ++d) {
    const int rule = node.rule[d];
    const int kind = rule & 15;
    if (kind == D2Q9_RULE_STREAM) continue;
    const d2q9_bc_elem& e = L.bc_elems[rule >> 4];
    const int o = INV[d];
    if (kind == D2Q9_RULE_BB) {
      g[d] = ld<S>(L, K, f, i, j, o);
    } else if (kind == D2Q9_RULE_VEL) {
      T rho_c = mom[c];
      T temp = ((((T)2 * K.w[o]) * rho_c) * ics2) *
               ((T)CX[o] * (T)e.vec[0] + (T)CY[o] * (T)e.vec[1]);
      g[d] = ld<S>(L, K, f, i, j, o) - temp;
    } else if (kind == D2Q9_RULE_PRES) {
      int64_t nb = c + (int64_t)e.normal[0] * L.pitch + e.normal[1];
      T vx = mom[L.plane + c], vy = mom[2 * L.plane + c];
      T ex = vx + (T)0.5 * (vx - mom[L.plane + nb]);
      T ey = vy + (T)0.5 * (vy - mom[2 * L.plane + nb]);
      T u2 = ex * ex + ey * ey;
      T cu = (T)CX[o] * ex + (T)CY[o] * ey;
      const T rb = (T)e.scalar;
      if constexpr (Codec<S>::dev) {
        // d = -d*[o] + (2 w rho_b (1 + ...) - 2 w); rho_b - 1 is formed in
        // double so that the small imposed density difference keeps its
        // digits in float
        const T drb = (T)(e.scalar - 1.0);
        T temp = ((T)2 * K.w[o]) *
                 (drb + rb * (((T)0.5 * ics4) * cu * cu - ((T)0.5 * ics2) * u2));
        g[d] = -ld<S>(L, K, f, i, j, o) + temp;
      } else {
        T temp = (((T)2 * K.w[o]) * rb) *
                 (((T)1 + ((T)0.5 * ics4) * cu * cu) - ((T)0.5 * ics2) * u2);
        g[d] = -ld<S>(L, K, f, i, j, o) + temp;
      }
    } else {  // D2Q9_RULE_ZG
      int ib = i + e.normal[0], jb = j + e.normal[1];
      int64_t nb = cell_index(L, ib, jb);
      uint8_t fnb = L.flags[nb];
      const T zero = Codec<S>::dev ? -K.w[d] : (T)0;
      g[d] = (fnb & D2Q9_CELL_FLUID) ? pull<S>(L, K, f, mom, ib, jb, nb, fnb, d) : zero;
    }
  }
}

// Workgroup b is dispatched to XCD b % 8; give each XCD a contiguous run of
// tiles so that tiles sharing cache lines meet in the same L2.
__device__ __forceinline__ int xcd_remap(int b, int n) {
  const int per = n >> 3, rem = n & 7;
  const int x = b & 7, r = b >> 3;
  return x * per + (x < rem ? x : rem) + r;
}

template <typename V>
__device__ __forceinline__ void store_nt(V v, V* p) {
  __builtin_nontemporal_store(v, p);
}

}  // namespace d2q9
