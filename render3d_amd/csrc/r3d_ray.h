// r3d_ray.h -- device helpers shared by the raycast forward and adjoint
// kernels: the pinned FP32 sampling sequence over the bricked pyramid.
#pragma once

#include "r3d_dmath.h"
#include "r3d_internal.h"

namespace r3d {

enum { MODE_FULL = 0, MODE_MIP = 1, MODE_DEPTH = 2, MODE_HEATMAP = 3, MODE_RAYDIR = 4, MODE_FLAT = 5 };
// Kernel kinds: FULL-mode views with the six-tap gradient, MIP, FLAT, the
// remaining modes, and FULL-mode views with the four-tap gradient (the
// reference's R3D_QUALITY_FAST pipelines, src/shaders/shaders.cmake:24-28).
enum { KIND_FULL = 0, KIND_MIP = 1, KIND_FLAT = 2, KIND_OTHER = 3, KIND_FAST = 4, NKINDS = 5 };
template <int KIND>
constexpr bool SHADED_KIND = KIND == KIND_FULL || KIND == KIND_FAST;
// The gradient's taps per shaded sample.
template <int KIND>
constexpr int KIND_TAPS = KIND == KIND_FAST ? 4 : 6;

constexpr float K255 = 1.0f / 255.0f;
constexpr uint32_t MAX_STEPS = 4096u;  // raycast.slang:120
constexpr uint32_t TILE = 8u;

struct V3 {
  float x, y, z;
};
__device__ __forceinline__ V3 mk(float x, float y, float z) { return V3{x, y, z}; }
__device__ __forceinline__ V3 vadd(V3 a, V3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ V3 vsub(V3 a, V3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ V3 vscale(V3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ float vdot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float vlen(V3 a) { return __builtin_sqrtf(vdot(a, a)); }
__device__ __forceinline__ V3 vnorm(V3 a) {
  const float l = vlen(a);
  return mk(a.x / l, a.y / l, a.z / l);
}

// A trilinear footprint: its four x-pairs of voxels as loaded (low byte at
// x0, high byte at x0 + 1) and the fractions. The bytes are taken apart
// only when the footprint is used, so a footprint can stay in flight.
struct TriFoot {
  uint32_t pr[4];
  float fx, fy, fz;
};

// WIDE: 64-bit byte offsets into the pyramid; otherwise (a pyramid below
// 4 GiB) 32-bit offsets on the scalar base, so every read is one
// base + 32-bit-offset load and the brick arithmetic stays in 32 bits.
template <bool WIDE>
struct SamplerT {
  const uint8_t *base;
  const LevelDev *lev;  // LDS
  uint32_t nlev;
  // measurement only: one bit per brick of the pyramid, set for every
  // footprint read (a compile-time null in the rendering kernels)
  uint32_t *touched = nullptr;

  using Foot = TriFoot;

  __device__ __forceinline__ static uint32_t pair_at(const uint8_t *q) {
    uint16_t v;
    __builtin_memcpy(&v, q, 2);
    return v;
  }

  __device__ __forceinline__ Foot load(uint32_t l, V3 p) const {
    const LevelDev &L = lev[l];
    const float tx = fclamp_ord(p.x * L.fx - 0.5f, 0.0f, L.hx);
    const float ty = fclamp_ord(p.y * L.fy - 0.5f, 0.0f, L.hy);
    const float tz = fclamp_ord(p.z * L.fz - 0.5f, 0.0f, L.hz);
    const uint32_t x0 = (uint32_t)tx, y0 = (uint32_t)ty, z0 = (uint32_t)tz;
    Foot f;
    f.fx = tx - (float)x0;
    f.fy = ty - (float)y0;
    f.fz = tz - (float)z0;
    const uint32_t b = ((z0 >> 4) * L.by + (y0 >> 4)) * L.bx + (x0 >> 4);
    const uint32_t in = ((z0 & 15u) * 17u + (y0 & 15u)) * 17u + (x0 & 15u);
    const uint8_t *q;
    if (WIDE) {
      const uint64_t off = ((uint64_t)L.off_hi << 32) | L.off_lo;
      q = base + off + (uint64_t)b * R3D_BRICK_BYTES + in;
      if (touched) {  // levels start 256-byte aligned, less than a brick past their predecessors' bricks
        const uint64_t gb = off / R3D_BRICK_BYTES + b;
        atomicOr(touched + (gb >> 5), 1u << (gb & 31u));
      }
    } else {
      q = base + (uint64_t)(L.off_lo + b * R3D_BRICK_BYTES + in);
    }
    f.pr[0] = pair_at(q);
    f.pr[1] = pair_at(q + 17);
    f.pr[2] = pair_at(q + 289);
    f.pr[3] = pair_at(q + 306);
    return f;
  }

  __device__ __forceinline__ static float lerp3(const Foot &f) {
    const float c00 = flerp((float)(f.pr[0] & 255u), (float)(f.pr[0] >> 8), f.fx);
    const float c10 = flerp((float)(f.pr[1] & 255u), (float)(f.pr[1] >> 8), f.fx);
    const float c01 = flerp((float)(f.pr[2] & 255u), (float)(f.pr[2] >> 8), f.fx);
    const float c11 = flerp((float)(f.pr[3] & 255u), (float)(f.pr[3] >> 8), f.fx);
    return flerp(flerp(c00, c10, f.fy), flerp(c01, c11, f.fy), f.fz);
  }

  __device__ __forceinline__ float tri(uint32_t l, V3 p) const { return lerp3(load(l, p)); }

  // Two samples at one lod with all four footprints loaded before any is
  // used (a scheduling barrier keeps the loads together): one memory round
  // trip for a pair of gradient taps.
  __device__ __forceinline__ void sample_pair(V3 pa, V3 pb, float lod, float *sa, float *sb) const {
    const uint32_t d = (uint32_t)lod;
    const float w = lod - (float)d;
    const bool blend = w > 0.0f && d + 1u < nlev;
    const uint32_t d1 = blend ? d + 1u : d;
    const Foot a0 = load(d, pa), a1 = load(d1, pa), b0 = load(d, pb), b1 = load(d1, pb);
    __builtin_amdgcn_sched_barrier(0);
    const float ra0 = lerp3(a0), ra1 = lerp3(a1), rb0 = lerp3(b0), rb1 = lerp3(b1);
    *sa = (blend ? flerp(ra0, ra1, w) : ra0) * K255;
    *sb = (blend ? flerp(rb0, rb1, w) : rb0) * K255;
  }

  // Both levels' footprints are read unconditionally (the second is level d
  // again when no blend is taken), so a sample costs one memory round trip
  // instead of two; the blend is then a select.
  __device__ __forceinline__ float sample(V3 p, float lod) const {
    const uint32_t d = (uint32_t)lod;
    const float w = lod - (float)d;
    const bool blend = w > 0.0f && d + 1u < nlev;
    const float s0 = tri(d, p), s1 = tri(blend ? d + 1u : d, p);
    const float s = blend ? flerp(s0, s1, w) : s0;
    return s * K255;
  }
};

using Sampler = SamplerT<true>;

// The occupancy byte of the 8^3 block holding p (clamped to the grid; the
// grid dims are at least 1).
__device__ __forceinline__ uint32_t occ_load(const VolDev &V, V3 p) {
  const float ox = __builtin_floorf(p.x * (float)V.ox);
  const float oy = __builtin_floorf(p.y * (float)V.oy);
  const float oz = __builtin_floorf(p.z * (float)V.oz);
  const uint32_t ix = (uint32_t)fclamp_ord(ox, 0.0f, (float)(V.ox - 1u));
  const uint32_t iy = (uint32_t)fclamp_ord(oy, 0.0f, (float)(V.oy - 1u));
  const uint32_t iz = (uint32_t)fclamp_ord(oz, 0.0f, (float)(V.oz - 1u));
  return V.occ[(iz * V.oy + iy) * V.ox + ix];  // fewer than 2^32 blocks in any volume that fits in memory
}

__device__ __forceinline__ float occ_sample(const VolDev &V, V3 p) { return (float)occ_load(V, p) * K255; }

__device__ __forceinline__ float4 tf_sample(const float4 *tf, float v) {
  const float t = fclamp_ord(v * 256.0f - 0.5f, 0.0f, 255.0f);
  const uint32_t i0 = (uint32_t)t;
  const uint32_t i1 = i0 + 1u < 256u ? i0 + 1u : 255u;
  const float f = t - (float)i0;
  const float4 a = tf[i0], b = tf[i1];
  return make_float4(flerp(a.x, b.x, f), flerp(a.y, b.y, f), flerp(a.z, b.z, f), flerp(a.w, b.w, f));
}

}  // namespace r3d
