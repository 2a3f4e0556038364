// r3d_ray.h -- device helpers shared by the raycast forward and adjoint
// kernels: the pinned FP32 sampling sequence over the bricked pyramid.
#pragma once

#include "r3d_dmath.h"
#include "r3d_internal.h"

namespace r3d {

enum { MODE_FULL = 0, MODE_MIP = 1, MODE_DEPTH = 2, MODE_HEATMAP = 3, MODE_RAYDIR = 4, MODE_FLAT = 5 };
enum { KIND_FULL = 0, KIND_MIP = 1, KIND_FLAT = 2, KIND_OTHER = 3, NKINDS = 4 };

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

struct Sampler {
  const uint8_t *base;
  const LevelDev *lev;  // LDS
  uint32_t nlev;

  __device__ __forceinline__ float tri(uint32_t l, V3 p) const {
    const LevelDev L = lev[l];
    const float tx = fclamp_sel(p.x * L.fx - 0.5f, 0.0f, L.fx - 1.0f);
    const float ty = fclamp_sel(p.y * L.fy - 0.5f, 0.0f, L.fy - 1.0f);
    const float tz = fclamp_sel(p.z * L.fz - 0.5f, 0.0f, L.fz - 1.0f);
    const uint32_t x0 = (uint32_t)tx, y0 = (uint32_t)ty, z0 = (uint32_t)tz;
    const float fx = tx - (float)x0, fy = ty - (float)y0, fz = tz - (float)z0;
    const uint32_t b = ((z0 >> 4) * L.by + (y0 >> 4)) * L.bx + (x0 >> 4);
    const uint64_t off = ((uint64_t)L.off_hi << 32) | L.off_lo;
    const uint8_t *q = base + off + (uint64_t)b * R3D_BRICK_BYTES +
                       (((z0 & 15u) * 17u + (y0 & 15u)) * 17u + (x0 & 15u));
    const float v000 = (float)q[0], v001 = (float)q[1], v010 = (float)q[17], v011 = (float)q[18];
    const float v100 = (float)q[289], v101 = (float)q[290], v110 = (float)q[306], v111 = (float)q[307];
    const float c00 = flerp(v000, v001, fx);
    const float c10 = flerp(v010, v011, fx);
    const float c01 = flerp(v100, v101, fx);
    const float c11 = flerp(v110, v111, fx);
    return flerp(flerp(c00, c10, fy), flerp(c01, c11, fy), fz);
  }

  __device__ __forceinline__ float sample(V3 p, float lod) const {
    const uint32_t d = (uint32_t)lod;
    const float w = lod - (float)d;
    float s = tri(d, p);
    if (w > 0.0f && d + 1u < nlev) s = flerp(s, tri(d + 1u, p), w);
    return s * K255;
  }
};

__device__ __forceinline__ float occ_sample(const VolDev &V, V3 p) {
  const float ox = __builtin_floorf(p.x * (float)V.ox);
  const float oy = __builtin_floorf(p.y * (float)V.oy);
  const float oz = __builtin_floorf(p.z * (float)V.oz);
  const uint32_t ix = (uint32_t)fclamp_sel(ox, 0.0f, (float)(V.ox - 1u));
  const uint32_t iy = (uint32_t)fclamp_sel(oy, 0.0f, (float)(V.oy - 1u));
  const uint32_t iz = (uint32_t)fclamp_sel(oz, 0.0f, (float)(V.oz - 1u));
  return (float)V.occ[((size_t)iz * V.oy + iy) * V.ox + ix] * K255;
}

__device__ __forceinline__ float4 tf_sample(const float4 *tf, float v) {
  const float t = fclamp_sel(v * 256.0f - 0.5f, 0.0f, 255.0f);
  const uint32_t i0 = (uint32_t)t;
  const uint32_t i1 = i0 + 1u < 256u ? i0 + 1u : 255u;
  const float f = t - (float)i0;
  const float4 a = tf[i0], b = tf[i1];
  return make_float4(flerp(a.x, b.x, f), flerp(a.y, b.y, f), flerp(a.z, b.z, f), flerp(a.w, b.w, f));
}

}  // namespace r3d
