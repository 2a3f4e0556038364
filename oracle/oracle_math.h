/* oracle_math.h -- TEST INFRASTRUCTURE ONLY (the oracle). Nothing in the
 * product links this file; only tests/, the smoke check and the CPU-baseline
 * leg of bench.py load the oracle library, as the checker.
 *
 * The pinned FP32 operation sequences for the transcendental functions the
 * raycaster needs. The reference shader calls the GPU's hardware log2/exp2/pow
 * (src/shaders/raycast.slang:776-779, 906, 993), whose ulp behaviour is
 * vendor-defined; this restatement pins one definition so a host program and a
 * HIP kernel that follow it produce identical bits:
 *
 *   log2f: x = 2^e * m, m in [1,2); if m > 1.41421356 then m *= 0.5, e += 1;
 *          f = m - 1; p = Horner(L9..L0, f) with fmaf; return (float)e + f*p.
 *          Subnormal x is pre-scaled by 2^23. x == 0 -> -inf, x < 0 or NaN
 *          -> NaN, +inf -> +inf.
 *   exp2f: x < -125 -> 0; x >= 127.5 -> +inf; NaN -> NaN; else
 *          k = floorf(x + 0.5); r = x - k; q = Horner(E6..E0, r) with fmaf;
 *          return q * 2^k (2^k built from its exponent bits).
 *   powf(b, y) for b in [0,1], y > 0 (the opacity correction): b <= 0 -> 0,
 *          b >= 1 -> 1, else exp2f(y * log2f(b)).
 *   pow24(x) = ((x^2)^2)^2 ... : x2=x*x, x4=x2*x2, x8=x4*x4, x16=x8*x8,
 *          x24 = x16*x8 (the specular exponent of raycast.slang:993).
 *
 * The coefficients are a Chebyshev-node least-squares fit (degree 9 for
 * log2(1+f)/f on [sqrt(1/2)-1, sqrt(2)-1], degree 6 for 2^r on [-1/2, 1/2]),
 * rounded to float. Measured: log2f <= 1.6 ulp, exp2f <= 1 ulp before scaling.
 * Every multiply-add that is fused is written as an explicit fmaf, and the file
 * is compiled with -ffp-contract=off, so no other operation is fused.
 */
#ifndef R3D_ORACLE_MATH_H
#define R3D_ORACLE_MATH_H

#include <math.h>
#include <stdint.h>
#include <string.h>

static inline uint32_t om_bits(float x) { uint32_t u; memcpy(&u, &x, 4); return u; }
static inline float om_float(uint32_t u) { float x; memcpy(&x, &u, 4); return x; }

/* min/max/clamp with the select semantics used everywhere in the pinned
 * sequence: (a < b) ? a : b. With a NaN operand the second operand wins. */
static inline float om_min(float a, float b) { return a < b ? a : b; }
static inline float om_max(float a, float b) { return a > b ? a : b; }
static inline float om_clamp(float x, float lo, float hi) { return om_min(om_max(x, lo), hi); }
static inline float om_sat(float x) { return om_clamp(x, 0.0f, 1.0f); }
/* lerp pinned as fmaf(t, b - a, a) (mathx.h:60 is a + (b - a) * t; fused here). */
static inline float om_lerp(float a, float b, float t) { return fmaf(t, b - a, a); }

static const float OM_L[10] = {
    -0x1.c362c0p-4f, 0x1.7d9132p-3f, -0x1.87381ap-3f, 0x1.a2f85ep-3f, -0x1.eabd64p-3f,
    0x1.277e9ap-2f,  -0x1.715a76p-2f, 0x1.ec7094p-2f, -0x1.715470p-1f, 0x1.715476p+0f,
};
static const float OM_E[7] = {
    0x1.446c7ep-13f, 0x1.5f48c8p-10f, 0x1.3b29d8p-7f, 0x1.c6aeccp-5f,
    0x1.ebfbe0p-3f,  0x1.62e430p-1f,  0x1.000000p+0f,
};

static inline float om_log2f(float x) {
  if (x != x || x < 0.0f) return NAN;
  if (x == 0.0f) return -INFINITY;
  if (x == INFINITY) return INFINITY;
  int32_t e0 = 0;
  if (x < 0x1p-126f) { x = x * 0x1p23f; e0 = -23; }
  uint32_t u = om_bits(x);
  int32_t e = (int32_t)((u >> 23) & 0xffu) - 127 + e0;
  float m = om_float((u & 0x7fffffu) | 0x3f800000u);
  if (m > 1.41421356f) { m = m * 0.5f; e = e + 1; }
  float f = m - 1.0f;
  float p = OM_L[0];
  for (int i = 1; i < 10; i++) p = fmaf(p, f, OM_L[i]);
  return (float)e + f * p;
}

static inline float om_exp2f(float x) {
  if (x != x) return x;
  if (x < -125.0f) return 0.0f;
  if (x >= 127.5f) return INFINITY;
  float k = floorf(x + 0.5f);
  float r = x - k;
  float q = OM_E[0];
  for (int i = 1; i < 7; i++) q = fmaf(q, r, OM_E[i]);
  int32_t ki = (int32_t)k;
  return q * om_float((uint32_t)(ki + 127) << 23);
}

static inline float om_pow_unit(float b, float y) {
  if (!(b > 0.0f)) return 0.0f;
  if (b >= 1.0f) return 1.0f;
  return om_exp2f(y * om_log2f(b));
}

static inline float om_pow24(float x) {
  float x2 = x * x, x4 = x2 * x2, x8 = x4 * x4, x16 = x8 * x8;
  return x16 * x8;
}

/* sin/cos of an angle given in turns (1 turn = 2*pi), used by the synthetic
 * volume generator: q = floor(4u + 0.5), a = 2*pi*(u - q/4) in [-pi/4, pi/4],
 * Taylor polynomials to degree 9 / 8 evaluated with fmaf. */
static inline void om_sincos_turns(float u, float *s, float *c) {
  float q = floorf(u * 4.0f + 0.5f);
  float t = u - q * 0.25f;
  float a = t * 6.28318530717958647692f;
  float a2 = a * a;
  float ps = fmaf(fmaf(fmaf(fmaf(2.7557319e-6f, a2, -1.98412698e-4f), a2, 8.33333333e-3f), a2,
                       -1.66666667e-1f), a2, 1.0f) * a;
  float pc = fmaf(fmaf(fmaf(fmaf(2.48015873e-5f, a2, -1.38888889e-3f), a2, 4.16666667e-2f), a2,
                       -0.5f), a2, 1.0f);
  int32_t qi = ((int32_t)q) & 3;
  float sv, cv;
  switch (qi) {
  case 0: sv = ps; cv = pc; break;
  case 1: sv = pc; cv = -ps; break;
  case 2: sv = -ps; cv = -pc; break;
  default: sv = -pc; cv = ps; break;
  }
  *s = sv;
  *c = cv;
}

/* hash12: common.slang:57-61, exact u32 arithmetic. */
static inline float om_hash12(uint32_t px, uint32_t py, uint32_t frame) {
  uint32_t h = px * 374761393u + py * 668265263u + frame * 2246822519u;
  h = (h ^ (h >> 13)) * 1274126177u;
  return (float)(h >> 8) * (1.0f / 16777216.0f);
}

#endif /* R3D_ORACLE_MATH_H */
