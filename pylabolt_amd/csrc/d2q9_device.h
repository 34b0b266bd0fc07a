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
// Random bits for rounding value `b` (float bits) of direction k of cell
// `salt`: a function of the data only, so a run is reproducible bit for
// bit, while the rounding of a slowly changing field is decorrelated from
// step to step.  The cell's key is hashed once per cell (cell_key, from
// full-rate 24-bit multiplies: a 32-bit multiply is quarter rate); per
// value one xor-add with the direction's salt (v_xad_u32) and one 24-bit
// multiply spread the low 24 bits of (value ^ key) + salt into the upper
// bits of the product, which the conversions below use: bits 16..31 for
// bfloat16, bits 19..31 for half.
__device__ __forceinline__ uint32_t cell_key(uint32_t salt) {
  uint32_t x = __umul24(salt, 0x9e3779u) ^ (salt >> 13);
  x = __umul24(x ^ (x >> 16), 0x85ebcau);
  return x ^ (x >> 15);
}
__device__ __forceinline__ uint32_t round_bits(uint32_t b, int k, uint32_t key) {
  const uint32_t x = (b ^ key) + (uint32_t)k * 0x632be5abu;
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
                                            float& fx, float& fy, float& px, float& py) {
  // pairs of opposite directions first: fewer dependent adds
  const float s13 = g[1] + g[3], s24 = g[2] + g[4], s57 = g[5] + g[7], s68 = g[6] + g[8];
  const float d13 = g[1] - g[3], d24 = g[2] - g[4], d57 = g[5] - g[7], d68 = g[6] - g[8];
  drho = (g[0] + s13) + (s24 + (s57 + s68));
  rho = 1.0f + drho;
  // c_x: +1 for 1,5,8; -1 for 3,6,7.  c_y: +1 for 2,5,6; -1 for 4,7,8.
  const float sx = d13 + (d57 - d68);
  const float sy = d24 + (d57 + d68);
  const float inv = __builtin_amdgcn_rcpf(rho + K.fmin);
  if constexpr (FORCING == D2Q9_FORCING_NONE) {
    // no force acts: gravity is zero without a forcing model
    // (base/force_operator.py:58-69), and a force field needs one
    fx = 0.0f;
    fy = 0.0f;
    px = sx;
    py = sy;
  } else {
    if (L.force_source == D2Q9_FSRC_FIELD) {
      fx = ffx;
      fy = ffy;
    } else {
      fx = rho * K.gx;
      fy = rho * K.gy;
    }
    px = __builtin_fmaf(0.5f, fx, sx);
    py = __builtin_fmaf(0.5f, fy, sy);
  }
  ux = px * inv;
  uy = py * inv;
}

template <typename T, int FORCING, bool DEV>
__device__ __forceinline__ void moments(const T (&g)[9], const d2q9_lattice& L,
                                        const Consts<T>& K, T ffx, T ffy,
                                        T& drho, T& rho, T& ux, T& uy, T& fx,
                                        T& fy, T& px, T& py) {
  if constexpr (std::is_same<T, float>::value && DEV) {
    moments_f32<FORCING>(g, L, K, ffx, ffy, drho, rho, ux, uy, fx, fy, px, py);
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
  if constexpr (FORCING == D2Q9_FORCING_NONE) {
    fx = (T)0;
    fy = (T)0;
  } else if (L.force_source == D2Q9_FSRC_FIELD) {
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
  if constexpr (FORCING != D2Q9_FORCING_NONE) {
    ux = sx * inv + ((T)0.5 * fx) * inv;
    uy = sy * inv + ((T)0.5 * fy) * inv;
  } else {
    // gravity is zero when forcing is None (base/force_operator.py:58-69),
    // so the force term adds exactly zero (and a force field needs a Guo
    // forcing)
    ux = sx * inv;
    uy = sy * inv;
  }
  // momentum sums (with half the force when forced); only the float
  // collisions read them
  px = sx + (T)0.5 * fx;
  py = sy + (T)0.5 * fy;
}

// Float collisions (deviation form) over the four opposite-direction
// pairs.  The equilibrium is written so that its mass and momentum equal
// the cell's up to the rounding of per-cell values: a constant factor of
// its own rounding (fl(3 w) with 3 (2 w1 + 4 w5) != 1 in binary, a weight
// sum != 1, fl(om w)) would put the same relative error into every update
// of a conserved quantity, which the collision does not relax away: it
// accumulates over the run (measured: FP32 cylinder 8192x2048 after 50,000
// steps, mean flow off by 1.7e-3).  So
//   geq_k + geq_kb = 2 w (drho + rho t_k),  t_k = 0.5 ics4 cu^2 - 0.5 ics2 u^2,
//   geq_k - geq_kb = 2 m (c_k . P),  P = sum_k c_k g_k (+ F/2) = rho u,
//   m1 = fl(ics2 w1) (axis pairs), m5 = (1 - 2 m1) / 4 (diagonal pairs),
//     so that 2 m1 + 4 m5 = 1 exactly and the momentum is P itself,
//   geq_0 = drho - sum_{k>0} geq_k  (mass drho),
// and the same for the Guo source (its shared parts sum to minus the rest
// direction's, its momentum is (1 - om/2) F through m).  omega multiplies
// per-cell values (rho, drho), never a weight.  In BGK the momentum term
// om m (c . P) is one product with the factors om m1 and om m5 =
// (om - 2 om m1) / 4, whose axis and diagonal shares again sum to om
// exactly: rounding om P first and m times that second had biased the
// momentum each step (FP32 cylinder after 50,000 steps: mean flow 1.26e-4
// slow; with one rounding 5e-6).
struct FloatEq {
  float hics4, hu2, m[2];
};

__device__ __forceinline__ FloatEq float_eq(const Consts<float>& K, float ux, float uy) {
  FloatEq E;
  E.hics4 = 0.5f * K.ics4;
  E.hu2 = (0.5f * K.ics2) * __builtin_fmaf(ux, ux, uy * uy);
  E.m[0] = K.ics2 * K.w[1];
  E.m[1] = 0.25f * __builtin_fmaf(-2.0f, E.m[0], 1.0f);
  return E;
}

// BGK: g_k <- (1 - om) g_k + om geq_k (+ (1 - om/2) Fbar_k).
template <int FORCING>
__device__ __forceinline__ void collide_bgk_pairs(float (&g)[9], const Consts<float>& K,
                                                  float drho, float rho, float ux, float uy,
                                                  float fx, float fy, float px, float py) {
  // fused multiply-adds are written out (never left to the compiler), so
  // every kernel build computes the same values
  constexpr bool forced = FORCING != D2Q9_FORCING_NONE;
  constexpr bool second = FORCING == D2Q9_GUO_SECOND_ORDER;
  const float om = K.om;
  const float omc = 1.0f - om;
  const FloatEq E = float_eq(K, ux, uy);
  const float cf = __builtin_fmaf(-0.5f, om, 1.0f);
  const float rom = rho * om, dom = drho * om;
  // om m1 and om m5, with 2 om m1 + 4 om m5 = om exactly (om - 2 om m1 is
  // exact: it lies within a factor of two of 2 om m1)
  const float om1 = E.m[0] * om;
  const float oms[2] = {om1, 0.25f * __builtin_fmaf(-2.0f, om1, om)};
  float cfx = 0.0f, cfy = 0.0f, uf = 0.0f;
  if constexpr (forced) {
    cfx = cf * fx;
    cfy = cf * fy;
  }
  if constexpr (second) uf = K.ics2 * __builtin_fmaf(ux, fx, uy * fy);
  constexpr int PK[4] = {1, 2, 5, 6};
  constexpr int PB[4] = {3, 4, 7, 8};
  float shared;
#pragma unroll
  for (int n = 0; n < 4; ++n) {
    const int k = PK[n], kb = PB[n];
    const float w = K.w[k], m = E.m[n >> 1];
    const float cu = cdot<float>(k, ux, uy);
    const float t = __builtin_fmaf(E.hics4 * cu, cu, -E.hu2);
    float A = w * __builtin_fmaf(rom, t, dom);
    float B = oms[n >> 1] * cdot<float>(k, px, py);
    if constexpr (forced) B = __builtin_fmaf(m, cdot<float>(k, cfx, cfy), B);
    if constexpr (second)
      A = __builtin_fmaf(cf * w, __builtin_fmaf(K.ics4 * cu, cdot<float>(k, fx, fy), -uf), A);
    shared = n ? shared + A : A;
    g[k] = __builtin_fmaf(omc, g[k], A + B);
    g[kb] = __builtin_fmaf(omc, g[kb], A - B);
  }
  g[0] = __builtin_fmaf(omc, g[0], __builtin_fmaf(-2.0f, shared, dom));
}

// BGK collision in place: g <- (1 - om) g + om geq (+ (1 - om/2) Fbar).
template <typename T, int FORCING, bool DEV>
__device__ __forceinline__ void collide_bgk(T (&g)[9], const Consts<T>& K,
                                            T drho, T rho, T ux, T uy, T fx, T fy,
                                            T px, T py) {
  if constexpr (std::is_same<T, float>::value && DEV) {
    collide_bgk_pairs<FORCING>(g, K, drho, rho, ux, uy, fx, fy, px, py);
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

// MRT with the reference's rates S = (1,1,1,1,1,1,1,s7,s8), s7 = s8 =
// omega (base/collision_operator.py:158-162).  The operator
//   f* = feq + M^-1 [(I - S) M (f - feq) + (I - S/2) M Fbar]
// has, with unit rates on rows 0..6, only the two stress rows left in the
// first term, and M^-1 = M^T diag(1/|row|^2) with |row 7|^2 = |row 8|^2 = 4:
//   f*_k = feq_k + 0.5 Fbar_k + sigma_k (1 - s) / 4 (p + 0.5 p^F),
// p = p_xx = (e1 + e3) - (e2 + e4) of e = f - feq (row 7) for k = 1..4,
// p_xy = (e5 + e7) - (e6 + e8) (row 8) for k = 5..8, p^F the same moment of
// Fbar, sigma_k = M[7][k] or M[8][k] (+1 for 1, 3, 5, 7; -1 for 2, 4, 6, 8);
// f*_0 = feq_0 + 0.5 Fbar_0.
//
// Float form (deviation storage), with the equilibrium and source of the
// float BGK above (mass and momentum exact up to per-cell rounding).  The
// stress moments of the equilibrium and the source have closed forms,
//   p_xx(geq) = w1 rho ics4 (ux^2 - uy^2),  p_xy(geq) = 4 w5 rho ics4 ux uy,
//   p_xx(Fbar) = 2 w1 ics4 (ux fx - uy fy), p_xy(Fbar) = 4 w5 ics4 (ux fy + uy fx)
// (second-order Guo; the linear source has none).  The same operator as
// the FP64 path, with other rounding.
template <int FORCING>
__device__ __forceinline__ void collide_mrt_pairs(float (&g)[9], const Consts<float>& K,
                                                  float drho, float rho, float ux, float uy,
                                                  float fx, float fy, float px, float py) {
  constexpr bool forced = FORCING != D2Q9_FORCING_NONE;
  constexpr bool second = FORCING == D2Q9_GUO_SECOND_ORDER;
  const float ics4 = K.ics4;
  const FloatEq E = float_eq(K, ux, uy);
  const float w1 = K.w[1], w5 = K.w[5];
  const float rics4 = rho * ics4;
  // stress moments of g - geq (+ half of Fbar's), times (1 - s) / 4
  float pxx = ((g[1] + g[3]) - (g[2] + g[4])) -
              (w1 * rics4) * __builtin_fmaf(ux, ux, -(uy * uy));
  float pxy = ((g[5] + g[7]) - (g[6] + g[8])) - ((4.0f * w5) * rics4) * (ux * uy);
  if constexpr (second) {
    pxx = __builtin_fmaf(w1 * ics4, __builtin_fmaf(ux, fx, -(uy * fy)), pxx);
    pxy = __builtin_fmaf((2.0f * w5) * ics4, __builtin_fmaf(ux, fy, uy * fx), pxy);
  }
  const float r7 = (0.25f * (1.0f - K.rates[7])) * pxx;
  const float r8 = (0.25f * (1.0f - K.rates[8])) * pxy;
  float uf = 0.0f;
  if constexpr (second) uf = K.ics2 * __builtin_fmaf(ux, fx, uy * fy);
  // The post-collision momentum is P + F/2 = sum c g + F.  Added to the
  // finished sum, the force is a nearly constant increment to a value that
  // changes little from step to step, and round-to-nearest then errs the
  // same way every step: in a channel starting from rest the mean flow
  // drifts (FP32, 4096x1024 MRT + Guo after 20,000 steps: 2.5e-4 below
  // FP64).  Added to g5 first, a diagonal deviation about a twelfth of the
  // sum with a finer ulp, the same rounding is far smaller (2e-5 in a
  // float emulation of that channel).
  float qx = px, qy = py;
  if constexpr (forced) {
    const float d13 = g[1] - g[3], d24 = g[2] - g[4], d68 = g[6] - g[8];
    qx = d13 + (((g[5] + fx) - g[7]) - d68);
    qy = d24 + (((g[5] + fy) - g[7]) + d68);
  }
  constexpr int PK[4] = {1, 2, 5, 6};
  constexpr int PB[4] = {3, 4, 7, 8};
  float shared;
#pragma unroll
  for (int n = 0; n < 4; ++n) {
    const int k = PK[n], kb = PB[n];
    const float w = K.w[k], m = E.m[n >> 1];
    const float cu = cdot<float>(k, ux, uy);
    const float t = __builtin_fmaf(E.hics4 * cu, cu, -E.hu2);
    float A = w * __builtin_fmaf(rho, t, drho);
    const float B = m * cdot<float>(k, qx, qy);
    if constexpr (second)
      A = __builtin_fmaf(0.5f * w, __builtin_fmaf(ics4 * cu, cdot<float>(k, fx, fy), -uf), A);
    shared = n ? shared + A : A;
    const float r = n < 2 ? r7 : r8;
    const float Ar = (n & 1) ? A - r : A + r;
    g[k] = Ar + B;
    g[kb] = Ar - B;
  }
  g[0] = __builtin_fmaf(-2.0f, shared, drho);
}

// MRT, FP64 storage, in the operation order of the oracle's mrt_collide
// (the reduced form above, with the stress moments of feq and Fbar in
// closed form):
//   q_k = feq_k (+ 0.5 Fbar_k),
//   a7 = ((f1 + f3) - (f2 + f4)) - ((w1 ics4) rho) (ux ux - uy uy)
//        (+ (w1 ics4) (ux fx - uy fy) with second-order Guo),
//   a8 = ((f5 + f7) - (f6 + f8)) - (((4 w5) ics4) rho) (ux uy)
//        (+ ((2 w5) ics4) (ux fy + uy fx) with second-order Guo),
//   r = ((1 - s) * 0.25) a,  f*_k = q_k +- r (f*_0 = q_0).
// The populations are consumed by the two stress sums; each output is then
// one equilibrium (and source) evaluation, so nothing of the nine is held.
template <typename T, int FORCING, bool DEV>
__device__ __forceinline__ void collide_mrt(T (&g)[9], const Consts<T>& K,
                                            T drho, T rho, T ux, T uy, T fx, T fy) {
  static_assert(!DEV, "deviation storage forms use collide_mrt_pairs");
  constexpr bool forced = FORCING != D2Q9_FORCING_NONE;
  const T ics2 = K.ics2, ics4 = K.ics4;
  const T w1 = K.w[1], w5 = K.w[5];
  T a7 = ((g[1] + g[3]) - (g[2] + g[4])) - ((w1 * ics4) * rho) * (ux * ux - uy * uy);
  T a8 = ((g[5] + g[7]) - (g[6] + g[8])) - ((((T)4 * w5) * ics4) * rho) * (ux * uy);
  if constexpr (FORCING == D2Q9_GUO_SECOND_ORDER) {
    a7 = a7 + (w1 * ics4) * (ux * fx - uy * fy);
    a8 = a8 + (((T)2 * w5) * ics4) * (ux * fy + uy * fx);
  }
  const T r7 = (((T)1 - K.rates[7]) * (T)0.25) * a7;
  const T r8 = (((T)1 - K.rates[8]) * (T)0.25) * a8;
  const T u2 = ux * ux + uy * uy;
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    T q = feq_k<T>(k, rho, ux, uy, u2, K.w[k], ics2, ics4);
    if constexpr (forced)
      q = q + (T)0.5 * guo_k<T, FORCING>(k, ux, uy, fx, fy, K.w[k], ics2, ics4);
    if (k == 0) g[0] = q;
    else if (k <= 4) g[k] = (k & 1) ? q + r7 : q - r7;
    else g[k] = (k & 1) ? q + r8 : q - r8;
  }
  (void)drho;
}

template <typename T, int MODEL, int FORCING, bool DEV>
__device__ __forceinline__ void collide(T (&g)[9], const Consts<T>& K, T drho,
                                        T rho, T ux, T uy, T fx, T fy, T px, T py) {
  if constexpr (MODEL == D2Q9_MRT) {
    if constexpr (std::is_same<T, float>::value && DEV)
      collide_mrt_pairs<FORCING>(g, K, drho, rho, ux, uy, fx, fy, px, py);
    else
      collide_mrt<T, FORCING, DEV>(g, K, drho, rho, ux, uy, fx, fy);
  } else
    collide_bgk<T, FORCING, DEV>(g, K, drho, rho, ux, uy, fx, fy, px, py);
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
