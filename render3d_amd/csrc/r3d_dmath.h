// r3d_dmath.h -- device FP32 routines with a pinned operation sequence.
//
// The reference shader calls the GPU's own log2/exp2/pow
// (src/shaders/raycast.slang:776-779, 906, 993), whose ulp behaviour is
// vendor-defined. The library pins one definition so its host transcription
// and these kernels agree bit for bit:
//
//   log2f: x = 2^e * m, m in [1,2); if m > 1.41421356 then m *= 0.5, e += 1;
//          f = m - 1; p = Horner(L9..L0, f) with fma; return (float)e + f*p.
//          Subnormal x is pre-scaled by 2^23; x == 0 -> -inf; x < 0, NaN -> NaN.
//   exp2f: x < -125 -> 0; x >= 127.5 -> +inf; else k = floor(x + 0.5),
//          r = x - k, q = Horner(E6..E0, r) with fma, return q * 2^k.
//   pow_unit(b, y), b in [0,1]: b <= 0 -> 0, b >= 1 -> 1, else exp2(y*log2 b).
//   pow24(x): x2 = x*x, x4, x8, x16 by squaring, x16 * x8.
//   sincos_turns(u): q = floor(4u + 0.5), a = 2 pi (u - q/4), Taylor to a^9.
//
// Coefficients: Chebyshev-node least-squares fits (degree 9 for log2(1+f)/f
// on [sqrt(1/2)-1, sqrt(2)-1], degree 6 for 2^r on [-1/2,1/2]) rounded to
// float. Everything else is plain IEEE FP32 (+ - * / sqrt are correctly
// rounded on gfx950 and on the host); the library is compiled with
// -ffp-contract=off so only the explicit fmaf calls are fused.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace r3d {

__device__ __forceinline__ float fmin_sel(float a, float b) { return a < b ? a : b; }
__device__ __forceinline__ float fmax_sel(float a, float b) { return a > b ? a : b; }
__device__ __forceinline__ float fclamp_sel(float x, float lo, float hi) {
  return fmin_sel(fmax_sel(x, lo), hi);
}
// fclamp_sel for lo <= hi in one v_med3_f32: the median of (x, lo, hi) is
// the select form's result for every x (NaN included: both give lo; checked
// bit for bit on the GPU over ordinary values, +-0, +-inf, NaN and random
// bit patterns).
__device__ __forceinline__ float fclamp_ord(float x, float lo, float hi) {
  return __builtin_amdgcn_fmed3f(x, lo, hi);
}
__device__ __forceinline__ float fsat(float x) { return fclamp_ord(x, 0.0f, 1.0f); }
// The u8 store of clamp(v, 0, 255) rounded half away from zero (roundf),
// written into byte b of w. With c = 0.49999997f (the float below 0.5),
// floor(v + c), the sum rounded to nearest even, is roundf(v) for every
// v >= 0 (a tie k + 0.5 sums to within half an ulp of k + 1 and rounds up to
// it; c < 0.5 keeps 0.5 - 2^-25 below 1), and at most 0 for v < 0.
// v_cvt_pk_u8_f32 rounds to nearest even and saturates to [0, 255] (NaN to
// 0; checked on this GPU over 4.2 M inputs), so on that integer it is exactly
// the clamp, and it inserts the byte in the same instruction. Identical to
// (uint8_t)roundf(clamp(v, 0, 255)) for every float (checked exhaustively
// over [-1, 65536]; beyond, both saturate).
__device__ __forceinline__ uint32_t round_u8_into(float v, uint32_t b, uint32_t w) {
  return __builtin_amdgcn_cvt_pk_u8_f32(__builtin_floorf(v + 0.49999997f), b, w);
}
// lerp pinned as fma(t, b - a, a)
__device__ __forceinline__ float flerp(float a, float b, float t) {
  return __builtin_fmaf(t, b - a, a);
}

// log2_pinned's arithmetic for a positive, finite, normal x (e0: the
// exponent bias of a pre-scaled subnormal).
__device__ __forceinline__ float log2_core(float x, int32_t e0 = 0) {
  uint32_t u = __float_as_uint(x);
  int32_t e = (int32_t)((u >> 23) & 0xffu) - 127 + e0;
  float m = __uint_as_float((u & 0x7fffffu) | 0x3f800000u);
  if (m > 1.41421356f) {
    m = m * 0.5f;
    e = e + 1;
  }
  float f = m - 1.0f;
  float p = -0x1.c362c0p-4f;
  p = __builtin_fmaf(p, f, 0x1.7d9132p-3f);
  p = __builtin_fmaf(p, f, -0x1.87381ap-3f);
  p = __builtin_fmaf(p, f, 0x1.a2f85ep-3f);
  p = __builtin_fmaf(p, f, -0x1.eabd64p-3f);
  p = __builtin_fmaf(p, f, 0x1.277e9ap-2f);
  p = __builtin_fmaf(p, f, -0x1.715a76p-2f);
  p = __builtin_fmaf(p, f, 0x1.ec7094p-2f);
  p = __builtin_fmaf(p, f, -0x1.715470p-1f);
  p = __builtin_fmaf(p, f, 0x1.715476p+0f);
  return (float)e + f * p;
}

__device__ __forceinline__ float log2_pinned(float x) {
  if (x != x || x < 0.0f) return __builtin_nanf("");
  if (x == 0.0f) return -__builtin_inff();
  if (x == __builtin_inff()) return x;
  const bool sub = x < 0x1p-126f;
  return log2_core(sub ? x * 0x1p23f : x, sub ? -23 : 0);
}

// log2_pinned for x >= 1 (NaN excluded): only +inf needs its own result,
// taken with a select instead of a branch.
__device__ __forceinline__ float log2_ge1(float x) {
  const float r = log2_core(x);
  return x < __builtin_inff() ? r : x;
}

// exp2_pinned's arithmetic for -125 <= x < 127.5.
__device__ __forceinline__ float exp2_core(float x) {
  float k = __builtin_floorf(x + 0.5f);
  float r = x - k;
  float q = 0x1.446c7ep-13f;
  q = __builtin_fmaf(q, r, 0x1.5f48c8p-10f);
  q = __builtin_fmaf(q, r, 0x1.3b29d8p-7f);
  q = __builtin_fmaf(q, r, 0x1.c6aeccp-5f);
  q = __builtin_fmaf(q, r, 0x1.ebfbe0p-3f);
  q = __builtin_fmaf(q, r, 0x1.62e430p-1f);
  q = __builtin_fmaf(q, r, 0x1.000000p+0f);
  int32_t ki = (int32_t)k;
  return q * __uint_as_float((uint32_t)(ki + 127) << 23);
}

__device__ __forceinline__ float exp2_pinned(float x) {
  if (x != x) return x;
  if (x < -125.0f) return 0.0f;
  if (x >= 127.5f) return __builtin_inff();
  return exp2_core(x);
}

__device__ __forceinline__ float pow_unit(float b, float y) {
  if (!(b > 0.0f)) return 0.0f;
  if (b >= 1.0f) return 1.0f;
  // 0 < b < 1: log2_pinned's special cases reduce to the subnormal scaling
  const bool sub = b < 0x1p-126f;
  return exp2_pinned(y * log2_core(sub ? b * 0x1p23f : b, sub ? -23 : 0));
}

__device__ __forceinline__ float pow24(float x) {
  float x2 = x * x, x4 = x2 * x2, x8 = x4 * x4, x16 = x8 * x8;
  return x16 * x8;
}

__device__ __forceinline__ void sincos_turns(float u, float *s, float *c) {
  float q = __builtin_floorf(u * 4.0f + 0.5f);
  float t = u - q * 0.25f;
  float a = t * 6.28318530717958647692f;
  float a2 = a * a;
  float ps = __builtin_fmaf(
                 __builtin_fmaf(__builtin_fmaf(__builtin_fmaf(2.7557319e-6f, a2, -1.98412698e-4f), a2,
                                               8.33333333e-3f),
                                a2, -1.66666667e-1f),
                 a2, 1.0f) *
             a;
  float pc = __builtin_fmaf(
      __builtin_fmaf(__builtin_fmaf(__builtin_fmaf(2.48015873e-5f, a2, -1.38888889e-3f), a2,
                                    4.16666667e-2f),
                     a2, -0.5f),
      a2, 1.0f);
  int32_t qi = ((int32_t)q) & 3;
  float sv, cv;
  if (qi == 0) {
    sv = ps;
    cv = pc;
  } else if (qi == 1) {
    sv = pc;
    cv = -ps;
  } else if (qi == 2) {
    sv = -ps;
    cv = -pc;
  } else {
    sv = -pc;
    cv = ps;
  }
  *s = sv;
  *c = cv;
}

// hash12 of src/shaders/common.slang:57-61 (exact u32 arithmetic).
__device__ __forceinline__ float hash12(uint32_t px, uint32_t py, uint32_t frame) {
  uint32_t h = px * 374761393u + py * 668265263u + frame * 2246822519u;
  h = (h ^ (h >> 13)) * 1274126177u;
  return (float)(h >> 8) * (1.0f / 16777216.0f);
}

}  // namespace r3d
