// r3d_raycast.hip -- cube-mode raycaster for gfx950.
//
// One lane per ray; a workgroup renders 8x8, 16x4, 16x8 or 16x16-pixel
// tiles, either one tile per workgroup or persistently from a tile queue (the
// shape and the persistence are chosen per plan by r3d_hip_raycast_plan_tune).
// Every view of a call renders in one launch per mode class (tile -> view
// through a prefix table). Per lane the march is src/shaders/raycast.slang main() with
// R3D_MODE 0 (lines 505-1039): ray setup, ray-cone LOD, trilinear + linear-mip
// sample, low cut, dilated-occupancy skip, TF lookup, opacity correction,
// 6-tap gradient Blinn-Phong, front-to-back compositing, early termination.
//
// The texture unit is replaced by a pinned FP32 sequence (the hardware's
// sub-texel rounding is not reproducible); a host transcription that follows
// the same sequence matches this kernel bit for bit:
//   tri(l, p):  t = clamp(p*dim - 0.5, 0, dim-1), i0 = (int)t, f = t - i0;
//               c00 = lerp(v000, v001, fx) ... lerp over x, then y, then z,
//               on the raw u8 values as floats
//   sample:     d = (int)lod, w = lod - d; s = tri(d); if (w > 0 && d+1 < nlev)
//               s = lerp(s, tri(d+1), w); v = s * (1/255)
//   lerp(a,b,t) = fma(t, b - a, a); dot products unfused; normalize = v / |v|
// The volume is the bricked pyramid of r3d_volume.hip: the 8 taps of a
// trilinear footprint are fixed offsets from one brick address, so a sample
// costs one address computation and 8 byte loads; the TF table and the level
// table sit in LDS.
#include <stdlib.h>
#include <string.h>

#include <new>

#include <type_traits>

#include "r3d_ray.h"

namespace r3d {

// Bricks-LOD volume (R3D_MODE 3 with the LOD flag): the reference's page
// table and 128^3 slots, slot s at atlas + s * 2 MiB.
struct BricksDev {
  const uint32_t *page;
  const uint8_t *atlas;
  uint32_t abpa;
  uint32_t npage, nslots;  // an entry past the table or a slot past the atlas reads as absent
};

struct RayArgs {
  VolDev vol;
  BricksDev bk;
  const r3d_frame_params *views;  // this launch's views
  const uint32_t *tile_start;     // nviews + 1 prefix sums of tiles
  uint32_t nviews;
  const float *tf;
  float *img;
  uint8_t *img8;
  uint32_t img_w, img_h;
  float cutoff;
  uint32_t *partials;  // [workgroup][8] counter sums of this launch, or NULL
  uint32_t *px_steps;
  uint32_t ntiles;     // tiles of this launch
  uint32_t *queue;     // tile counter of a persistent launch, or NULL
  uint32_t *touched;   // measurement: one bit per brick a footprint reads, or NULL
  const uint32_t *order;  // launch order: the i-th workgroup (or queue entry) renders tile order[i]; NULL: tile i
  uint32_t tier2;      // k_raycast_tiers: ranks below tier2 take two lanes per ray
};

// Samples a ray issues per memory round trip, per kernel kind.
#ifndef R3D_RAY_BATCH_FULL
#define R3D_RAY_BATCH_FULL 2
#endif
#ifndef R3D_RAY_BATCH_MIP
#define R3D_RAY_BATCH_MIP 3
#endif
#ifndef R3D_RAY_BATCH_FLAT
#define R3D_RAY_BATCH_FLAT 3
#endif
template <int KIND>
constexpr int RAY_BATCH = SHADED_KIND<KIND> ? R3D_RAY_BATCH_FULL
                          : KIND == KIND_MIP ? R3D_RAY_BATCH_MIP
                                             : R3D_RAY_BATCH_FLAT;

struct RayOut {
  float c[4];
  uint32_t steps, skips, shaded;
  uint32_t blends, blends_shaded;  // samples (and shaded samples) that blend two mip levels
};

// Wave-cooperative gradient taps (FULL): the lanes whose sample is shaded
// publish (p, e, aux) in LDS, every active lane evaluates one tap of one of
// them per round, and the results come back through LDS. A wave with k
// shaded lanes and n active lanes pays ceil(NTAP k / n) samples instead of
// NTAP. Each tap is the fetch the owner would make itself:
//   six taps (R3D_QUALITY_FULL): p +x, -x, +y, -y, +z, -z at radius e;
//   four taps (R3D_QUALITY_FAST): p + e (+,-,-), (-,-,+), (-,+,-), (+,+,+),
//   the owner publishing the tetrahedron radius (raycast.slang:925-953).
// buf holds 64 x 5 + 384 floats.
constexpr uint32_t COOP_FLOATS = 64u * 5u + 384u;
constexpr float IR3 = 0.57735026919f;     // 1/sqrt(3) (raycast.slang:926)
constexpr float RENORM = 0.86602540378f;  // sqrt(3)/2 (raycast.slang:927)

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ uint32_t lanes_below(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// fetch(q, aux) evaluates one tap at position q; aux is the owner's fifth
// published word (the cube path's lod, the bricks path's slot).
template <int NTAP, class Fetch>
__device__ __forceinline__ void coop_taps(const Fetch &fetch, float *buf, bool need, V3 p, float e, float aux,
                                          float g[NTAP]) {
  const uint64_t act = __ballot(1);
  const uint64_t m = __ballot(need);
  if (m == 0) return;
  const uint32_t k = (uint32_t)__popcll(m), nact = (uint32_t)__popcll(act);
  const uint32_t rn = lanes_below(m), ra = lanes_below(act);
  float *pub = buf, *res = buf + 320;
  if (need) {
    pub[rn * 5u + 0u] = p.x;
    pub[rn * 5u + 1u] = p.y;
    pub[rn * 5u + 2u] = p.z;
    pub[rn * 5u + 3u] = e;
    pub[rn * 5u + 4u] = aux;
  }
  wave_sync();
  const uint32_t njobs = (uint32_t)NTAP * k;
  for (uint32_t base = 0; base < njobs; base += nact) {
    const uint32_t job = base + ra;
    if (job < njobs) {
      const uint32_t o = job / (uint32_t)NTAP, tap = job - (uint32_t)NTAP * o;
      const float qx = pub[o * 5u + 0u], qy = pub[o * 5u + 1u], qz = pub[o * 5u + 2u];
      const float qe = pub[o * 5u + 3u], ql = pub[o * 5u + 4u];
      V3 q;
      if (NTAP == 6) {
        const bool neg = (tap & 1u) != 0u;
        const float sh = tap < 2u ? qx : tap < 4u ? qy : qz;
        const float moved = neg ? sh - qe : sh + qe;
        q = mk(tap < 2u ? moved : qx, tap - 2u < 2u ? moved : qy, tap >= 4u ? moved : qz);
      } else {  // x + -e is x - e
        const bool px = tap == 0u || tap == 3u, py = tap >= 2u, pz = tap == 1u || tap == 3u;
        q = mk(px ? qx + qe : qx - qe, py ? qy + qe : qy - qe, pz ? qz + qe : qz - qe);
      }
      res[job] = fetch(q, ql);
    }
  }
  wave_sync();
  if (need) {
#pragma unroll
    for (int i = 0; i < NTAP; i++) g[i] = res[rn * (uint32_t)NTAP + (uint32_t)i];
  }
  wave_sync();  // the buffer is reused by the next step
}

// The tap radius a shaded lane publishes: e for the six taps, e / sqrt(3)
// for the tetrahedron (raycast.slang:941-945).
template <int NTAP>
__device__ __forceinline__ float tap_radius(float e) {
  return NTAP == 4 ? e * IR3 : e;
}

// The gradient from the taps of coop_taps: central differences, or the
// tetrahedral combination renormalised by sqrt(3)/2 (raycast.slang:952-953),
// left to right in FP32.
template <int NTAP>
__device__ __forceinline__ V3 tap_gradient(const float *q) {
  if (NTAP == 4)
    return mk((q[0] - q[1] - q[2] + q[3]) * RENORM, (-q[0] - q[1] + q[2] + q[3]) * RENORM,
              (-q[0] + q[1] - q[2] + q[3]) * RENORM);
  return mk(q[0] - q[1], q[2] - q[3], q[4] - q[5]);
}

// Gradient-modulated Blinn-Phong (raycast.slang:985-997): headlight
// l = -dir, key light normalize(0.4, 0.8, -0.4) rounded once, half vector
// normalize(l - dir); homogeneous regions (|g| <= 1e-4) stay unshaded.
__device__ __forceinline__ void blinn_phong(V3 g, V3 lvec, V3 hvec, float &cr, float &cg, float &cb) {
  const V3 key = mk(0x1.a20bd8p-2f, 0x1.a20bd8p-1f, -0x1.a20bd8p-2f);
  const float gm = vlen(g);
  if (gm > 1e-4f) {
    const V3 n = mk(g.x / gm, g.y / gm, g.z / gm);
    const float diff = 0.7f * __builtin_fabsf(vdot(n, lvec)) + 0.3f * __builtin_fabsf(vdot(n, key));
    const float spec = pow24(__builtin_fabsf(vdot(n, hvec))) * 0.3f;
    const float lit = 0.25f + 0.75f * diff;
    const float w = fsat(gm * 6.0f);
    const float m = flerp(1.0f, lit, w);
    const float sw = spec * w;
    cr = cr * m + sw;
    cg = cg * m + sw;
    cb = cb * m + sw;
  }
}

// Whether a sample of value v can contribute: opq holds one bit per TF
// interval [i, i + 1], set unless both entries have zero alpha. A sample in
// a zero interval gets alpha lerp(0, 0, f) = 0, so its opacity
// 1 - pow(1 - 0, k) is 0 and it adds nothing; i0 is tf_sample's index.
__device__ __forceinline__ bool tf_visible(const uint32_t *opq, float v) {
  const uint32_t i0 = (uint32_t)fclamp_ord(v * 256.0f - 0.5f, 0.0f, 255.0f);
  return ((opq[i0 >> 5] >> (i0 & 31u)) & 1u) != 0u;
}

// A cube-mode ray up to its march loop: ray setup, model transform, box hit,
// cone, step control and jitter (raycast.slang:469-495, 677-728, 743-748).
struct RaySetup {
  V3 dir, ro, lvec, hvec;
  float tn, tfar, t0, voxel, fdim, max_lod, cone, dt, span_step, bs, occ_gate, opac_k, sgx, sgy, sgz;
  bool ortho, dim_p2;
};

// Fills G and returns true when the ray marches; otherwise R holds its pixel
// (the RAYDIR colour, or the background of a ray that misses the box).
template <int KIND>
__device__ __forceinline__ bool ray_setup(const VolDev &V, const r3d_frame_params &pc, uint32_t mode, uint32_t px,
                                          uint32_t py, RaySetup &G, RayOut &R) {
  const V3 fwd = mk(pc.cam_forward[0], pc.cam_forward[1], pc.cam_forward[2]);
  const V3 right = mk(pc.cam_right[0], pc.cam_right[1], pc.cam_right[2]);
  const V3 up = mk(pc.cam_up[0], pc.cam_up[1], pc.cam_up[2]);
  const bool ortho = (pc.view_flags & 1u) != 0u;
  const float u = ((float)px + 0.5f) / (float)pc.viewport[0];
  const float vv = ((float)py + 0.5f) / (float)pc.viewport[1];
  const float ndx = u * 2.0f - 1.0f, ndy = 1.0f - vv * 2.0f;
  V3 dir = ortho ? vnorm(fwd) : vnorm(vadd(vadd(fwd, vscale(right, ndx)), vscale(up, ndy)));
  if (KIND == KIND_OTHER && mode == MODE_RAYDIR) {
    R.c[0] = dir.x * 0.5f + 0.5f;
    R.c[1] = dir.y * 0.5f + 0.5f;
    R.c[2] = dir.z * 0.5f + 0.5f;
    R.c[3] = 1.0f;
    return false;
  }
  G.ortho = ortho;
  G.voxel = 1.0f / (float)V.maxdim;
  // with a power-of-two largest dim, voxel is exact and x / voxel == x * maxdim
  G.dim_p2 = (V.maxdim & (V.maxdim - 1u)) == 0u;
  G.fdim = (float)V.maxdim;
  G.max_lod = fmin_sel(pc.max_mip, (float)(V.nlev - 1u));
  V3 org = mk(pc.cam_origin[0], pc.cam_origin[1], pc.cam_origin[2]);
  if (ortho) org = vadd(vadd(org, vscale(right, ndx)), vscale(up, ndy));
  // common.slang:45-52 with ext = (1,1,1), center 0.5
  const V3 ctr = mk(0.5f, 0.5f, 0.5f);
  const V3 c0 = mk(pc.vol_r0[0], pc.vol_r1[0], pc.vol_r2[0]);
  const V3 c1 = mk(pc.vol_r0[1], pc.vol_r1[1], pc.vol_r2[1]);
  const V3 c2 = mk(pc.vol_r0[2], pc.vol_r1[2], pc.vol_r2[2]);
  const V3 q = vsub(vsub(org, mk(pc.vol_tx, pc.vol_ty, pc.vol_tz)), ctr);
  const V3 ro = vadd(mk(vdot(c0, q), vdot(c1, q), vdot(c2, q)), ctr);
  dir = mk(vdot(c0, dir), vdot(c1, dir), vdot(c2, dir));
  // box_hit against [0,1]^3 (raycast.slang:489-495)
  const V3 inv = mk(1.0f / dir.x, 1.0f / dir.y, 1.0f / dir.z);
  const V3 t0 = mk((0.0f - ro.x) * inv.x, (0.0f - ro.y) * inv.y, (0.0f - ro.z) * inv.z);
  const V3 t1 = mk((1.0f - ro.x) * inv.x, (1.0f - ro.y) * inv.y, (1.0f - ro.z) * inv.z);
  const float hx = fmax_sel(fmax_sel(fmin_sel(t0.x, t1.x), fmin_sel(t0.y, t1.y)), fmin_sel(t0.z, t1.z));
  const float hy = fmin_sel(fmin_sel(fmax_sel(t0.x, t1.x), fmax_sel(t0.y, t1.y)), fmax_sel(t0.z, t1.z));
  const float tn = fmax_sel(hx, 0.0f), tfar = fmin_sel(hy, 1e30f);
  if (tn >= tfar) {
    R.c[0] = 0.02f;
    R.c[1] = 0.02f;
    R.c[2] = 0.03f;
    R.c[3] = 1.0f;
    return false;
  }
  G.dir = dir;
  G.ro = ro;
  G.tn = tn;
  G.tfar = tfar;
  G.cone = 2.0f * vlen(up) / (float)pc.viewport[1];
  G.dt = pc.step_voxels * G.voxel;
  G.t0 = tn + hash12(px, py, pc.frame_index) * G.dt;
  G.span_step = fmax_sel((tfar - tn) / ((float)MAX_STEPS * G.dt), 1.0f);
  G.bs = 8.0f * G.voxel;
  G.occ_gate = fmax_sel(pc.skip_gate, 0.5f / 255.0f);
  G.opac_k = pc.density * pc.step_voxels;
  G.lvec = vscale(dir, -1.0f);
  G.hvec = vnorm(vsub(G.lvec, dir));
  G.sgx = dir.x > 0.0f ? 1.0f : 0.0f;
  G.sgy = dir.y > 0.0f ? 1.0f : 0.0f;
  G.sgz = dir.z > 0.0f ? 1.0f : 0.0f;
  return true;
}

// The lod and lod step of a sample at t (raycast.slang:767-780); fv >= 1 and
// max_lod in [0, 15], so log2 and exp2 need none of their special cases.
__device__ __forceinline__ void lod_at(const RaySetup &G, float lod_bias, float tt, float &lod, float &lod_step) {
  const float foot = G.ortho ? G.cone : tt * G.cone;
  const float fv = G.dim_p2 ? foot * G.fdim : foot / G.voxel;
  lod = fclamp_ord(log2_ge1(fmax_sel(fv, 1.0f)) + lod_bias, 0.0f, G.max_lod);
  lod_step = fmax_sel(exp2_core(lod), G.span_step);
}

// Where the empty-space skip of a sample at (t, p) resumes: the next 8^3
// block face along the ray plus 1% of a block (raycast.slang:865-868).
__device__ __forceinline__ float skip_target(const RaySetup &G, V3 p, float t) {
  const float bs = G.bs;
  const float nx = (__builtin_floorf(p.x / bs) + G.sgx) * bs;
  const float ny = (__builtin_floorf(p.y / bs) + G.sgy) * bs;
  const float nz = (__builtin_floorf(p.z / bs) + G.sgz) * bs;
  const float ex = __builtin_fabsf(G.dir.x) > 1e-8f ? (nx - G.ro.x) / G.dir.x : 1e30f;
  const float ey = __builtin_fabsf(G.dir.y) > 1e-8f ? (ny - G.ro.y) / G.dir.y : 1e30f;
  const float ez = __builtin_fabsf(G.dir.z) > 1e-8f ? (nz - G.ro.z) / G.dir.z : 1e30f;
  return fmax_sel(fmin_sel(fmin_sel(ex, ey), ez), t) + bs * 0.01f;
}

// The pixel of a finished march (raycast.slang:1023-1038).
template <int KIND>
__device__ __forceinline__ void ray_finish(RayOut &R, uint32_t mode, const r3d_frame_params &pc, uint32_t steps,
                                           float ar, float ag, float ab, float aa, float mip_max, float first_hit) {
  const bool composite = mode == MODE_FULL || mode == MODE_FLAT;
  if (KIND == KIND_OTHER && mode == MODE_HEATMAP) {
    const float x = (float)steps / 1024.0f;
    R.c[0] = x;
    R.c[1] = 1.0f - __builtin_fabsf(x - 0.5f) * 2.0f;
    R.c[2] = 1.0f - x;
  } else if (KIND == KIND_OTHER && mode == MODE_DEPTH) {
    const float d = first_hit < 0.0f ? 0.0f : 1.0f - fsat(first_hit / 2.0f);
    R.c[0] = d;
    R.c[1] = d;
    R.c[2] = d;
  } else if (composite) {
    const float ta = 1.0f - aa;
    R.c[0] = ar + ta * 0.02f;
    R.c[1] = ag + ta * 0.02f;
    R.c[2] = ab + ta * 0.03f;
  } else {
    const float m = mip_max * pc.density;
    R.c[0] = m;
    R.c[1] = m;
    R.c[2] = m;
  }
  R.c[3] = 1.0f;
}

template <int KIND>
constexpr uint32_t KIND_MODE = SHADED_KIND<KIND> ? MODE_FULL : KIND == KIND_MIP ? MODE_MIP : MODE_FLAT;

// The per-ray march (MEASURE: also count blended samples, for the
// measurement instances only). KIND fixes the compositing / shading branches at compile
// time, as the reference builds one pipeline per mode (src/shaders/shaders.cmake:17-30).
template <int KIND, bool MEASURE, class SamplerK>
__device__ __forceinline__ RayOut march(const SamplerK &S, const VolDev &V, const float4 *tf, const uint32_t *opq,
                                        const r3d_frame_params &pc, float cutoff, uint32_t px, uint32_t py,
                                        float *coop) {
  RayOut R;
  R.steps = R.skips = R.shaded = R.blends = R.blends_shaded = 0;
  const uint32_t mode = KIND == KIND_OTHER ? pc.mode : KIND_MODE<KIND>;
  RaySetup G;
  if (!ray_setup<KIND>(V, pc, mode, px, py, G, R)) return R;
  const V3 dir = G.dir, ro = G.ro, lvec = G.lvec, hvec = G.hvec;
  const float tfar = G.tfar, voxel = G.voxel, dt = G.dt;
  const float occ_gate = G.occ_gate, opac_k = G.opac_k;
  float t = G.t0;
  float mip_max = 0.0f, first_hit = -1.0f;
  float ar = 0.0f, ag = 0.0f, ab = 0.0f, aa = 0.0f;
  uint32_t steps = 0, skips = 0, shaded_n = 0, blends = 0, blends_sh = 0;
  const bool composite = mode == MODE_FULL || mode == MODE_FLAT;
  // One sample's position, lod and reads (both levels' footprints and the
  // occupancy byte). Reads are clamped into the volume, so a sample issued
  // past the end of the ray is always valid to read and simply dropped.
  struct Pending {
    float lod, lod_step;
    V3 p;
    TriFoot f0, f1;
    uint32_t occ;
  };
  // view fields the loop reads, held in registers (a read of the view in the
  // loop would wait behind the sample reads in flight)
  const float lod_bias = pc.lod_bias, threshold = pc.threshold;
  const auto issue = [&](float tt) {
    Pending q;
    lod_at(G, lod_bias, tt, q.lod, q.lod_step);
    q.p = vadd(ro, vscale(dir, tt));
    const uint32_t d = (uint32_t)q.lod;
    const bool blend = q.lod - (float)d > 0.0f && d + 1u < S.nlev;
    q.f0 = S.load(d, q.p);
    q.f1 = S.load(blend ? d + 1u : d, q.p);
    q.occ = occ_load(V, q.p);
    return q;
  };
  // One march step on the sample in cur (issued at the current t): 0 ends
  // the ray, 1 advanced t by dt * lod_step, 2 skipped to another t.
  const auto step = [&](const Pending &cur) -> int {
    steps++;
    const float lod = cur.lod, lod_step = cur.lod_step;
    const V3 p = cur.p;
    const float t_next = t + dt * lod_step;
    const uint32_t d = (uint32_t)lod;
    const float w = lod - (float)d;
    const float s0 = S.lerp3(cur.f0), s1 = S.lerp3(cur.f1);
    const bool blend = w > 0.0f && d + 1u < S.nlev;
    if (MEASURE) blends += blend ? 1u : 0u;
    float v = (blend ? flerp(s0, s1, w) : s0) * K255;
    const float occ = (float)cur.occ * K255;
    const bool cut = v < threshold;
    if (cut) v = 0.0f;
    if ((v < occ_gate) & (occ < occ_gate)) {  // raycast.slang:866-869
      t = skip_target(G, p, t);
      skips++;
      return 2;
    }
    if (v > mip_max) mip_max = v;
    if (mode == MODE_MIP && mip_max >= 0.999f) return 0;
    if (KIND == KIND_OTHER && first_hit < 0.0f && v > 0.05f) first_hit = t;
    if (SHADED_KIND<KIND>) {
      // every active lane takes part in the taps, so the TF and opacity are
      // evaluated for all of them (a cut sample simply does not use them);
      // when no lane's sample can contribute, the wave skips all of it
      if (__ballot(!cut && tf_visible(opq, v)) != 0) {
        constexpr int NTAP = KIND_TAPS<KIND>;
        const float4 s = tf_sample(tf, v);
        const float a = 1.0f - pow_unit(1.0f - s.w, opac_k * lod_step);
        const bool lit_here = !cut && a > 0.003f;
        float gt[NTAP];
        coop_taps<NTAP>([&](V3 q, float ql) { return S.sample(q, ql); }, coop, lit_here, p,
                        tap_radius<NTAP>(voxel * lod_step), lod, gt);
        if (lit_here) {
          float cr = s.x, cg = s.y, cb = s.z;
          shaded_n++;
          if (MEASURE) blends_sh += blend ? 1u : 0u;
          blinn_phong(tap_gradient<NTAP>(gt), lvec, hvec, cr, cg, cb);
          const float wa = (1.0f - aa) * a;
          ar = ar + wa * cr;
          ag = ag + wa * cg;
          ab = ab + wa * cb;
          aa = aa + wa;
          if (aa >= cutoff) return 0;
        }
      }
    } else if (composite && !cut && tf_visible(opq, v)) {
      const float4 s = tf_sample(tf, v);
      const float a = 1.0f - pow_unit(1.0f - s.w, opac_k * lod_step);
      if (a > 0.003f) {
        const float wa = (1.0f - aa) * a;
        ar = ar + wa * s.x;
        ag = ag + wa * s.y;
        ab = ab + wa * s.z;
        aa = aa + wa;
        if (aa >= cutoff) return 0;
      }
    }
    t = t_next;
    return 1;
  };
  // BATCH consecutive samples are issued together (one memory round trip
  // for all); each is used only if the ones before it neither ended the ray
  // nor skipped (a skip issues a new batch from the new t).
  constexpr int BATCH = MEASURE ? 1 : RAY_BATCH<KIND>;  // measuring: no sample is read and dropped
  bool live = true;
  while (live && t < tfar && steps < MAX_STEPS) {
    Pending q[BATCH];
    float tq = t;
#pragma unroll
    for (int i = 0; i < BATCH; i++) {
      q[i] = issue(tq);
      tq = tq + dt * q[i].lod_step;
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int i = 0; i < BATCH; i++) {
      if (i > 0 && !(t < tfar && steps < MAX_STEPS)) break;
      const int r = step(q[i]);
      if (r == 0) live = false;
      if (r != 1) break;
    }
  }
  R.steps = steps;
  R.skips = skips;
  R.shaded = shaded_n;
  R.blends = blends;
  R.blends_shaded = blends_sh;
  ray_finish<KIND>(R, mode, pc, steps, ar, ag, ab, aa, mip_max, first_hit);
  return R;
}

// ---- lane-sharded march ------------------------------------------------------
// A long ray's march is a chain of dependent samples. For the costliest tiles
// a ray gets a pair of adjacent lanes: each iteration the pair takes the
// ray's next two samples (lane j the j-th: both lanes walk the step-control
// chain from the ray's t and keep their own position), reads and evaluates
// them side by side, and then folds them into the ray's state in sample order
// through pair broadcasts. Both lanes apply the same fold, so they hold the
// same state. A sample is the same computation as in march() and the fold
// runs march()'s operations in its order; a sample after one that ends the
// ray or skips is dropped, so image, counters and per-pixel steps are
// march()'s. The extra chain step and the fold cost about a fifth of a
// shaded sample's instructions and halve the longest rays' chains, which
// bound the frame; unshaded modes (MIP, FLAT) lose more than they gain, and
// four lanes per ray measured slower than two.
constexpr int LPR = 2;

// v from lane SRC of this lane's pair (a quad permutation).
template <int SRC>
__device__ __forceinline__ uint32_t pair_bcast(uint32_t v) {
  constexpr int ctrl = SRC | SRC << 2 | (SRC + 2) << 4 | (SRC + 2) << 6;
  return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, ctrl, 0xf, 0xf, false);
}
template <int SRC>
__device__ __forceinline__ float pair_bcast(float v) {
  return __uint_as_float(pair_bcast<SRC>(__float_as_uint(v)));
}

template <int KIND, class SamplerK>
__device__ __forceinline__ RayOut march_pair(const SamplerK &S, const VolDev &V, const float4 *tf,
                                             const uint32_t *opq, const r3d_frame_params &pc, float cutoff,
                                             uint32_t px, uint32_t py, uint32_t j, float *coop) {
  static_assert(SHADED_KIND<KIND>, "the lane-sharded march renders the shaded modes");
  RayOut R;
  R.steps = R.skips = R.shaded = R.blends = R.blends_shaded = 0;
  constexpr uint32_t mode = KIND_MODE<KIND>;
  RaySetup G;
  if (!ray_setup<KIND>(V, pc, mode, px, py, G, R)) return R;
  const float lod_bias = pc.lod_bias, threshold = pc.threshold;
  float T = G.t0;  // the ray's next sample
  float ar = 0.0f, ag = 0.0f, ab = 0.0f, aa = 0.0f;
  uint32_t steps = 0, skips = 0, shaded_n = 0;
  bool live = true;
  while (live && T < G.tfar && steps < MAX_STEPS) {
    // this lane's sample: the j-th of the chain from T
    float tj = T, lodj = 0.0f, stepj = 1.0f, tc = T;
#pragma unroll
    for (int i = 0; i < LPR; i++) {
      float lod, ls;
      lod_at(G, lod_bias, tc, lod, ls);
      if ((uint32_t)i == j) {
        tj = tc;
        lodj = lod;
        stepj = ls;
      }
      tc = tc + G.dt * ls;
    }
    const V3 p = vadd(G.ro, vscale(G.dir, tj));
    const uint32_t d = (uint32_t)lodj;
    const float w = lodj - (float)d;
    const bool blend = w > 0.0f && d + 1u < S.nlev;
    const TriFoot f0 = S.load(d, p), f1 = S.load(blend ? d + 1u : d, p);
    const uint32_t occb = occ_load(V, p);
    const bool valid = tj < G.tfar && steps + j < MAX_STEPS;
    const float s0 = S.lerp3(f0), s1 = S.lerp3(f1);
    float v = (blend ? flerp(s0, s1, w) : s0) * K255;
    const float occ = (float)occb * K255;
    const bool cut = v < threshold;
    if (cut) v = 0.0f;
    const bool skip = (v < G.occ_gate) & (occ < G.occ_gate);
    // the state after this sample: the skip target, or the next t
    const float tnext = skip ? skip_target(G, p, tj) : tj + G.dt * stepj;
    float a = 0.0f, cr = 0.0f, cg = 0.0f, cb = 0.0f;
    bool lit = false;
    const bool open = valid && !skip;
    if (__ballot(open && !cut && tf_visible(opq, v)) != 0) {
      constexpr int NTAP = KIND_TAPS<KIND>;
      const float4 s = tf_sample(tf, v);
      a = 1.0f - pow_unit(1.0f - s.w, G.opac_k * stepj);
      lit = open && !cut && a > 0.003f;
      float gt[NTAP];
      coop_taps<NTAP>([&](V3 q, float ql) { return S.sample(q, ql); }, coop, lit, p,
                      tap_radius<NTAP>(G.voxel * stepj), lodj, gt);
      if (lit) {
        cr = s.x;
        cg = s.y;
        cb = s.z;
        blinn_phong(tap_gradient<NTAP>(gt), G.lvec, G.hvec, cr, cg, cb);
      }
    }
    // fold the pair's samples into the ray's state, in sample order
    const uint32_t flags = (valid ? 1u : 0u) | (skip ? 2u : 0u) | (lit ? 4u : 0u);
    bool more = true;  // no sample of this iteration has ended the ray or skipped yet
    const auto fold = [&](auto src) {
      constexpr int SRC = decltype(src)::value;
      const uint32_t f = pair_bcast<SRC>(flags);
      const float tn = pair_bcast<SRC>(tnext);
      const float as = pair_bcast<SRC>(a), rs = pair_bcast<SRC>(cr), gs = pair_bcast<SRC>(cg),
                  bs = pair_bcast<SRC>(cb);
      if (!more) return;
      if (!(f & 1u)) {  // t reached tfar or the step budget: the loop ends
        live = false;
        more = false;
        return;
      }
      steps++;
      if (f & 2u) {  // skipped: the next iteration starts at the skip target
        T = tn;
        skips++;
        more = false;
        return;
      }
      if (f & 4u) {
        shaded_n++;
        const float wa = (1.0f - aa) * as;
        ar = ar + wa * rs;
        ag = ag + wa * gs;
        ab = ab + wa * bs;
        aa = aa + wa;
        if (aa >= cutoff) {
          live = false;
          more = false;
          return;
        }
      }
      T = tn;
    };
    fold(std::integral_constant<int, 0>{});
    fold(std::integral_constant<int, 1>{});
  }
  R.steps = steps;
  R.skips = skips;
  R.shaded = shaded_n;
  ray_finish<KIND>(R, mode, pc, steps, ar, ag, ab, aa, 0.0f, -1.0f);
  return R;
}

// ---- bricks-LOD mode -----------------------------------------------------
// raycast.slang main() with R3D_MODE 3 and BR_LOD_FLAG: page-table locate
// with a per-ray cache (lines 198-279), single-level fetch inside a 128^3
// slot, whole-brick skips, taps in brick-local voxels. The fetch is the
// pinned trilinear of the cube path: t = clamp(bl, 0.5, 127.5) - 0.5,
// i1 = min(i0 + 1, 127), lerp over x, y, z, times 1/255.
constexpr float BR_SLOT = 128.0f;
constexpr uint64_t BR_SLOT_VOX = 128ull * 128ull * 128ull;

struct BCtx {  // flags as full words: a bool's padding bytes were copied through scratch
  uint32_t ok;
  float maxv;
  V3 bl;
  uint32_t slot;
  float grad_vox, world_brick;
};

struct LodCache {
  float inv;  // 1 / scale of the cached level
  uint32_t valid;
  int desired;
  uint32_t scale, bx, by, bz;
  BCtx c;
};

// A slot footprint as loaded (bytes taken apart at use) and its fractions.
struct BFoot {
  uint32_t v[8];
  float fx, fy, fz;
};

__device__ __forceinline__ BFoot brick_load(const uint8_t *atlas, uint32_t slot, V3 bl) {
  const float tx = fclamp_ord(bl.x, 0.5f, 127.5f) - 0.5f;
  const float ty = fclamp_ord(bl.y, 0.5f, 127.5f) - 0.5f;
  const float tz = fclamp_ord(bl.z, 0.5f, 127.5f) - 0.5f;
  const uint32_t x0 = (uint32_t)tx, y0 = (uint32_t)ty, z0 = (uint32_t)tz;
  const uint32_t dx = x0 < 127u ? 1u : 0u, dy = y0 < 127u ? 128u : 0u, dz = z0 < 127u ? 16384u : 0u;
  BFoot f;
  f.fx = tx - (float)x0;
  f.fy = ty - (float)y0;
  f.fz = tz - (float)z0;
  const uint8_t *q = atlas + (uint64_t)slot * BR_SLOT_VOX + ((z0 * 128u + y0) * 128u + x0);
  f.v[0] = q[0];
  f.v[1] = q[dx];
  f.v[2] = q[dy];
  f.v[3] = q[dy + dx];
  f.v[4] = q[dz];
  f.v[5] = q[dz + dx];
  f.v[6] = q[dz + dy];
  f.v[7] = q[dz + dy + dx];
  return f;
}

__device__ __forceinline__ float brick_lerp(const BFoot &f) {
  const float c00 = flerp((float)f.v[0], (float)f.v[1], f.fx);
  const float c10 = flerp((float)f.v[2], (float)f.v[3], f.fx);
  const float c01 = flerp((float)f.v[4], (float)f.v[5], f.fx);
  const float c11 = flerp((float)f.v[6], (float)f.v[7], f.fx);
  return flerp(flerp(c00, c10, f.fy), flerp(c01, c11, f.fy), f.fz) * K255;
}

__device__ __forceinline__ float brick_fetch(const uint8_t *atlas, uint32_t slot, V3 bl) {
  return brick_lerp(brick_load(atlas, slot, bl));
}

__device__ __forceinline__ V3 base_voxel(const uint32_t *hdr, V3 p, float md) {
  const float sx = (float)hdr[1], sy = (float)hdr[2], sz = (float)hdr[3];
  return mk(fclamp_ord(p.x * md, 0.0f, fmax_sel(sx - 0.001f, 0.0f)), fclamp_ord(p.y * md, 0.0f, fmax_sel(sy - 0.001f, 0.0f)),
            fclamp_ord(p.z * md, 0.0f, fmax_sel(sz - 0.001f, 0.0f)));
}

// bricks_lod_locate_ex: desired level, +1, +2, then the coarsest. hdr is the
// page header in LDS; entries are read from the page table.
// scales_p2: every level's scale is a power of two (as built here), so
// x / scale == x * (1 / scale) exactly and the divisions become products.
__device__ __forceinline__ BCtx locate_ex(const BricksDev &B, const uint32_t *hdr, float md, V3 p, float lod,
                                          bool scales_p2, uint32_t *scale_out, float *inv_out, uint32_t *bx_o,
                                          uint32_t *by_o, uint32_t *bz_o) {
  BCtx c;
  c.ok = 0u;
  c.maxv = 0.0f;
  c.bl = mk(0.0f, 0.0f, 0.0f);
  c.slot = 0;
  c.grad_vox = 1.0f;
  *scale_out = 1u;
  *inv_out = 1.0f;
  *bx_o = *by_o = *bz_o = 0u;
  const uint32_t nlev = hdr[0];
  const V3 bv = base_voxel(hdr, p, md);
  c.world_brick = BR_SLOT * (float)(1u << (nlev - 1u)) / md;
  const int desired = (int)fclamp_ord(__builtin_floorf(lod), 0.0f, (float)(nlev - 1u));
  const int last = (int)nlev - 1;
  for (uint32_t step = 0; step < 4u; step++) {
    const int li = step < 3u ? desired + (int)step : last;
    if (li < 0 || li > last || (step == 3u && li <= desired + 2)) continue;
    const uint32_t h = 4u + (uint32_t)li * 4u;
    const uint32_t off = hdr[h], packed = hdr[h + 1u], scale = hdr[h + 2u];
    const uint32_t bdx = packed & 1023u, bdy = (packed >> 10) & 1023u, bdz = packed >> 20;
    const float fs = (float)scale, inv = 1.0f / fs;
    const V3 lv = scales_p2 ? mk(bv.x * inv, bv.y * inv, bv.z * inv) : mk(bv.x / fs, bv.y / fs, bv.z / fs);
    const uint32_t bcx = min((uint32_t)lv.x >> 7, bdx - 1u), bcy = min((uint32_t)lv.y >> 7, bdy - 1u),
                   bcz = min((uint32_t)lv.z >> 7, bdz - 1u);
    const uint32_t w = off + (bcz * bdy + bcy) * bdx + bcx;
    const uint32_t e = w < B.npage ? B.page[w] : 0xFFFFFFFFu;
    if (e == 0xFFFFFFFFu || (e & 0xFFFFFFu) >= B.nslots) continue;
    c.ok = 1u;
    c.maxv = (float)(e >> 24) / 255.0f;
    c.slot = e & 0xFFFFFFu;
    c.bl = mk(lv.x - (float)(bcx << 7), lv.y - (float)(bcy << 7), lv.z - (float)(bcz << 7));
    c.grad_vox = scales_p2 ? exp2_core(lod) * inv : exp2_core(lod) / fs;  // lod in [0, 14]
    c.world_brick = BR_SLOT * fs / md;
    *scale_out = scale;
    *inv_out = inv;
    *bx_o = bcx;
    *by_o = bcy;
    *bz_o = bcz;
    return c;
  }
  return c;
}

// bricks_lod_locate_cached: the walk is skipped while the sample stays in the
// cached brick at the same desired level (its level, grad_vox and brick size
// are then the cached ones).
// The cache's answer for a sample at p: true (and the located brick in c)
// when the sample stays in the cached brick at the same desired level.
__device__ __forceinline__ bool cache_probe(const uint32_t *hdr, float md, V3 p, float lod, bool scales_p2,
                                            const LodCache &lc, BCtx *c) {
  const uint32_t nlev = hdr[0];
  const int desired = (int)fclamp_ord(__builtin_floorf(lod), 0.0f, (float)(nlev - 1u));
  if (!(lc.valid && lc.desired == desired)) return false;
  const V3 bv = base_voxel(hdr, p, md);
  const float fs = (float)lc.scale, inv = lc.inv;
  const V3 lv = scales_p2 ? mk(bv.x * inv, bv.y * inv, bv.z * inv) : mk(bv.x / fs, bv.y / fs, bv.z / fs);
  const uint32_t bcx = (uint32_t)lv.x >> 7, bcy = (uint32_t)lv.y >> 7, bcz = (uint32_t)lv.z >> 7;
  if (!(bcx == lc.bx && bcy == lc.by && bcz == lc.bz)) return false;
  *c = lc.c;
  c->bl = mk(lv.x - (float)(bcx << 7), lv.y - (float)(bcy << 7), lv.z - (float)(bcz << 7));
  return true;
}

__device__ __forceinline__ BCtx locate_cached(const BricksDev &B, const uint32_t *hdr, float md, V3 p, float lod,
                                              bool scales_p2, LodCache &lc) {
  BCtx hit;
  if (cache_probe(hdr, md, p, lod, scales_p2, lc, &hit)) return hit;
  const uint32_t nlev = hdr[0];
  const int desired = (int)fclamp_ord(__builtin_floorf(lod), 0.0f, (float)(nlev - 1u));
  uint32_t scale, bx, by, bz;
  float inv;
  const BCtx c = locate_ex(B, hdr, md, p, lod, scales_p2, &scale, &inv, &bx, &by, &bz);
  lc.valid = c.ok;
  lc.desired = desired;
  lc.scale = scale;
  lc.inv = inv;
  lc.bx = bx;
  lc.by = by;
  lc.bz = bz;
  lc.c = c;
  return c;
}

template <int KIND>
__device__ __forceinline__ RayOut march_bricks(const BricksDev &B, const uint32_t *hdr, const float4 *tf,
                                               const uint32_t *opq, const r3d_frame_params &pc, float cutoff,
                                               uint32_t px, uint32_t py, float *coop) {
  RayOut R;
  R.steps = R.skips = R.shaded = R.blends = R.blends_shaded = 0;  // one level per sample: no blends
  const uint32_t mode = SHADED_KIND<KIND> ? MODE_FULL : KIND == KIND_MIP ? MODE_MIP : KIND == KIND_FLAT ? MODE_FLAT : pc.mode;
  const V3 fwd = mk(pc.cam_forward[0], pc.cam_forward[1], pc.cam_forward[2]);
  const V3 right = mk(pc.cam_right[0], pc.cam_right[1], pc.cam_right[2]);
  const V3 up = mk(pc.cam_up[0], pc.cam_up[1], pc.cam_up[2]);
  const bool ortho = (pc.view_flags & 1u) != 0u;
  const float u = ((float)px + 0.5f) / (float)pc.viewport[0];
  const float vv = ((float)py + 0.5f) / (float)pc.viewport[1];
  const float ndx = u * 2.0f - 1.0f, ndy = 1.0f - vv * 2.0f;
  V3 dir = ortho ? vnorm(fwd) : vnorm(vadd(vadd(fwd, vscale(right, ndx)), vscale(up, ndy)));
  if (KIND == KIND_OTHER && mode == MODE_RAYDIR) {
    R.c[0] = dir.x * 0.5f + 0.5f;
    R.c[1] = dir.y * 0.5f + 0.5f;
    R.c[2] = dir.z * 0.5f + 0.5f;
    R.c[3] = 1.0f;
    return R;
  }
  const float sx = (float)hdr[1], sy = (float)hdr[2], sz = (float)hdr[3];
  const uint32_t mdu = max(hdr[1], max(hdr[2], hdr[3]));
  const float md = (float)mdu;
  const float voxel = 1.0f / md;
  const bool md_p2 = (mdu & (mdu - 1u)) == 0u;  // then x / voxel == x * md
  // levels' scales, all powers of two in a table built here (wave-uniform)
  bool scales_p2 = true;
  for (uint32_t l = 0; l < hdr[0]; l++) {
    const uint32_t sc = hdr[6u + 4u * l];
    scales_p2 = scales_p2 && sc != 0u && (sc & (sc - 1u)) == 0u;
  }
  scales_p2 = __builtin_amdgcn_readfirstlane(scales_p2 ? 1u : 0u) != 0u;
  const float max_lod = (float)(hdr[0] - 1u);
  const float emax = fmax_sel(sx, fmax_sel(sy, sz));
  const V3 ext = mk(sx / emax, sy / emax, sz / emax);
  V3 org = mk(pc.cam_origin[0], pc.cam_origin[1], pc.cam_origin[2]);
  if (ortho) org = vadd(vadd(org, vscale(right, ndx)), vscale(up, ndy));
  const V3 ctr = vscale(ext, 0.5f);
  const V3 c0 = mk(pc.vol_r0[0], pc.vol_r1[0], pc.vol_r2[0]);
  const V3 c1 = mk(pc.vol_r0[1], pc.vol_r1[1], pc.vol_r2[1]);
  const V3 c2 = mk(pc.vol_r0[2], pc.vol_r1[2], pc.vol_r2[2]);
  const V3 q = vsub(vsub(org, mk(pc.vol_tx, pc.vol_ty, pc.vol_tz)), ctr);
  const V3 ro = vadd(mk(vdot(c0, q), vdot(c1, q), vdot(c2, q)), ctr);
  dir = mk(vdot(c0, dir), vdot(c1, dir), vdot(c2, dir));
  // raycast.slang:690-718: zoom-crop box or slab range
  V3 lo = mk(0.0f, 0.0f, 0.0f), hi = ext;
  float t_lo = 0.0f, t_hi = 1e30f;
  if ((pc.view_flags & 16u) != 0u) {
    const V3 clo = mk(pc.slab_x0 / md, pc.slab_y0 / md, pc.slab_z0 / md);
    const V3 chi = vadd(clo, mk(pc.slab_px / md, pc.slab_py / md, pc.slab_nx / md));
    lo = mk(fclamp_sel(clo.x, 0.0f, ext.x), fclamp_sel(clo.y, 0.0f, ext.y), fclamp_sel(clo.z, 0.0f, ext.z));
    hi = mk(fclamp_sel(chi.x, lo.x, ext.x), fclamp_sel(chi.y, lo.y, ext.y), fclamp_sel(chi.z, lo.z, ext.z));
  } else if (pc.slab_depth > 0u) {
    if ((pc.view_flags & 4u) != 0u) {
      t_lo = pc.slab_z0 / md;
      t_hi = (pc.slab_z0 + (float)pc.slab_depth) / md;
    } else {
      const uint32_t axc = (pc.view_flags >> 8) & 3u;
      const float a0 = fclamp_sel(pc.slab_z0 / md, 0.0f, axc == 1u ? ext.x : axc == 2u ? ext.y : ext.z);
      const float a1 = fclamp_sel((pc.slab_z0 + (float)pc.slab_depth) / md, a0, axc == 1u ? ext.x : axc == 2u ? ext.y : ext.z);
      if (axc == 1u) {
        lo.x = a0;
        hi.x = a1;
      } else if (axc == 2u) {
        lo.y = a0;
        hi.y = a1;
      } else {
        lo.z = a0;
        hi.z = a1;
      }
    }
  }
  const V3 inv = mk(1.0f / dir.x, 1.0f / dir.y, 1.0f / dir.z);
  const V3 t0 = mk((lo.x - ro.x) * inv.x, (lo.y - ro.y) * inv.y, (lo.z - ro.z) * inv.z);
  const V3 t1 = mk((hi.x - ro.x) * inv.x, (hi.y - ro.y) * inv.y, (hi.z - ro.z) * inv.z);
  const float hx = fmax_sel(fmax_sel(fmin_sel(t0.x, t1.x), fmin_sel(t0.y, t1.y)), fmin_sel(t0.z, t1.z));
  const float hy = fmin_sel(fmin_sel(fmax_sel(t0.x, t1.x), fmax_sel(t0.y, t1.y)), fmax_sel(t0.z, t1.z));
  const float tn = fmax_sel(hx, t_lo), tfar = fmin_sel(hy, t_hi);
  if (tn >= tfar) {
    R.c[0] = 0.02f;
    R.c[1] = 0.02f;
    R.c[2] = 0.03f;
    R.c[3] = 1.0f;
    return R;
  }
  const float cone = 2.0f * vlen(up) / (float)pc.viewport[1];
  const float dt = pc.step_voxels * voxel;
  float t = tn + hash12(px, py, pc.frame_index) * dt;
  float mip_max = 0.0f, first_hit = -1.0f;
  float ar = 0.0f, ag = 0.0f, ab = 0.0f, aa = 0.0f;
  uint32_t steps = 0, skips = 0, shaded_n = 0;
  const bool composite = mode == MODE_FULL || mode == MODE_FLAT;
  const bool shaded = mode == MODE_FULL;
  const uint32_t budget = shaded ? 192u : 256u;  // raycast.slang:743
  const float span_step = fmax_sel((tfar - tn) / ((float)budget * dt), 1.0f);
  const float occ_gate = fmax_sel(pc.skip_gate, 0.5f / 255.0f);
  const float bgate = fmax_sel(occ_gate, 4.5f / 255.0f);
  const float opac_k = pc.density * pc.step_voxels;
  const V3 lvec = vscale(dir, -1.0f);
  const V3 hvec = vnorm(vsub(lvec, dir));
  const float sgx = dir.x > 0.0f ? 1.0f : 0.0f, sgy = dir.y > 0.0f ? 1.0f : 0.0f,
              sgz = dir.z > 0.0f ? 1.0f : 0.0f;
  LodCache lc;
  lc.valid = 0u;
  lc.desired = -1;
  lc.scale = 1u;
  lc.inv = 1.0f;
  lc.bx = lc.by = lc.bz = 0u;
  lc.c = BCtx{false, 0.0f, mk(0.0f, 0.0f, 0.0f), 0u, 1.0f, 0.0f};
  // A sample: position, lod, located brick and its footprint's reads.
  struct BPend {
    float lod, lod_step;
    V3 p;
    BCtx bc;
    BFoot f;
  };
  const auto at_t = [&](float tt, BPend &q) {
    const float foot = ortho ? cone : tt * cone;
    const float fv = md_p2 ? foot * md : foot / voxel;
    // fv >= 1 and max_lod in [0, 14]: log2 and exp2 without their special cases
    q.lod = fclamp_ord(log2_ge1(fmax_sel(fv, 1.0f)) + pc.lod_bias, 0.0f, max_lod);
    q.lod_step = fmax_sel(exp2_core(q.lod), span_step);
    q.p = vadd(ro, vscale(dir, tt));
  };
  // One march step on a located, issued sample at the current t: 0 ends the
  // ray, 1 advanced t by dt * lod_step, 2 skipped.
  const auto step = [&](const BPend &q) -> int {
    steps++;
    const float lod_step = q.lod_step;
    const V3 p = q.p;
    const BCtx &bc = q.bc;
    float v = bc.ok ? brick_lerp(q.f) : 0.0f;
    const bool cut = v < pc.threshold;
    if (cut) v = 0.0f;
    // raycast.slang:834-868: whole-brick skips
    float bsz = 0.0f;
    if (mode == MODE_MIP && bc.ok && bc.maxv <= mip_max)
      bsz = bc.world_brick;
    else if (v < bgate && bc.maxv < bgate)
      bsz = bc.world_brick;
    if (bsz > 0.0f) {
      const float nx = (__builtin_floorf(p.x / bsz) + sgx) * bsz;
      const float ny = (__builtin_floorf(p.y / bsz) + sgy) * bsz;
      const float nz = (__builtin_floorf(p.z / bsz) + sgz) * bsz;
      const float ex = __builtin_fabsf(dir.x) > 1e-8f ? (nx - ro.x) / dir.x : 1e30f;
      const float ey = __builtin_fabsf(dir.y) > 1e-8f ? (ny - ro.y) / dir.y : 1e30f;
      const float ez = __builtin_fabsf(dir.z) > 1e-8f ? (nz - ro.z) / dir.z : 1e30f;
      t = fmax_sel(fmin_sel(fmin_sel(ex, ey), ez), t) + bsz * 0.01f;
      skips++;
      return 2;
    }
    if (v > mip_max) mip_max = v;
    if (mode == MODE_MIP && mip_max >= 0.999f) return 0;
    if (KIND == KIND_OTHER && first_hit < 0.0f && v > 0.05f) first_hit = t;
    if (SHADED_KIND<KIND>) {  // every active lane takes part in the taps (see march)
      if (__ballot(!cut && tf_visible(opq, v)) != 0) {
        constexpr int NTAP = KIND_TAPS<KIND>;
        const float4 s = tf_sample(tf, v);
        const float a = 1.0f - pow_unit(1.0f - s.w, opac_k * lod_step);
        const bool lit_here = !cut && a > 0.003f;
        float gt[NTAP];
        // taps reuse the located brick (raycast.slang:933-937, 964-973); x + -ev is x - ev
        coop_taps<NTAP>([&](V3 q, float slot) { return brick_fetch(B.atlas, __builtin_bit_cast(uint32_t, slot), q); },
                        coop, lit_here, bc.bl, tap_radius<NTAP>(bc.grad_vox), __builtin_bit_cast(float, bc.slot), gt);
        if (lit_here) {
          float cr = s.x, cg = s.y, cb = s.z;
          shaded_n++;
          blinn_phong(tap_gradient<NTAP>(gt), lvec, hvec, cr, cg, cb);
          const float wa = (1.0f - aa) * a;
          ar = ar + wa * cr;
          ag = ag + wa * cg;
          ab = ab + wa * cb;
          aa = aa + wa;
          if (aa >= cutoff) return 0;
        }
      }
    } else if (composite && !cut && tf_visible(opq, v)) {
      const float4 s = tf_sample(tf, v);
      const float a = 1.0f - pow_unit(1.0f - s.w, opac_k * lod_step);
      if (a > 0.003f) {
        const float wa = (1.0f - aa) * a;
        ar = ar + wa * s.x;
        ag = ag + wa * s.y;
        ab = ab + wa * s.z;
        aa = aa + wa;
        if (aa >= cutoff) return 0;
      }
    }
    t += dt * lod_step;
    return 1;
  };
  // The sample after a located one is located from the same cache when it
  // stays in the cached brick (pure arithmetic), so both footprints are read
  // in one memory round trip; otherwise it is located in the next iteration.
  // Reads of an unlocated sample go to slot 0 and are not used.
  while (t < tfar && steps < budget) {
    BPend a, b;
    at_t(t, a);
    a.bc = locate_cached(B, hdr, md, a.p, a.lod, scales_p2, lc);
    a.f = brick_load(B.atlas, a.bc.ok ? a.bc.slot : 0u, a.bc.bl);
    at_t(t + dt * a.lod_step, b);
    b.bc = BCtx{0u, 0.0f, mk(0.0f, 0.0f, 0.0f), 0u, 1.0f, 0.0f};  // slot 0 unless the probe hits
    const bool hit_b = cache_probe(hdr, md, b.p, b.lod, scales_p2, lc, &b.bc);
    b.f = brick_load(B.atlas, b.bc.slot, b.bc.bl);
    __builtin_amdgcn_sched_barrier(0);
    const int ra = step(a);
    if (ra == 0) break;
    if (ra == 2 || !hit_b || !(t < tfar && steps < budget)) continue;
    if (step(b) == 0) break;
  }
  R.steps = steps;
  R.skips = skips;
  R.shaded = shaded_n;
  if (KIND == KIND_OTHER && mode == MODE_HEATMAP) {
    const float x = (float)steps / 1024.0f;
    R.c[0] = x;
    R.c[1] = 1.0f - __builtin_fabsf(x - 0.5f) * 2.0f;
    R.c[2] = 1.0f - x;
  } else if (KIND == KIND_OTHER && mode == MODE_DEPTH) {
    const float d = first_hit < 0.0f ? 0.0f : 1.0f - fsat(first_hit / 2.0f);
    R.c[0] = d;
    R.c[1] = d;
    R.c[2] = d;
  } else if (composite) {
    const float ta = 1.0f - aa;
    R.c[0] = ar + ta * 0.02f;
    R.c[1] = ag + ta * 0.02f;
    R.c[2] = ab + ta * 0.03f;
  } else {
    const float m = mip_max * pc.density;
    R.c[0] = m;
    R.c[1] = m;
    R.c[2] = m;
  }
  R.c[3] = 1.0f;
  return R;
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Workgroup tile shapes in pixels, one lane per pixel.
constexpr int NSHAPES = 4;
constexpr uint32_t SHAPE_W[NSHAPES] = {8u, 16u, 16u, 16u};
constexpr uint32_t SHAPE_H[NSHAPES] = {8u, 4u, 8u, 16u};
constexpr uint32_t MAX_BLOCKS_PER_CU = 32u;

static int shape_index(uint32_t w, uint32_t h) {
  for (int i = 0; i < NSHAPES; i++)
    if (SHAPE_W[i] == w && SHAPE_H[i] == h) return i;
  return -1;
}

// One workgroup renders TW x TH-pixel tiles of the launch's views. With
// a.queue == NULL the workgroup renders tile blockIdx.x; otherwise it is
// persistent and takes tiles from the queue counter until they run out (the
// TF and level tables are then staged once per workgroup, not once per tile).
// Counter partials: two uint4 per workgroup (samples, skips, shaded, rays;
// blended samples, blended shaded samples, 0, 0). MEASURE (8x8 cube-mode
// measurement instances only) counts the blended samples and marks every
// brick a footprint reads in a.touched (when set), one bit per brick.
template <int KIND, int TW, int TH, bool BRICKS, bool MEASURE = false, bool WIDE = true>
__global__ __launch_bounds__(TW * TH) void k_raycast(RayArgs a) {
  constexpr uint32_t NT = TW * TH, NW = NT / 64u;
  __shared__ float4 s_tf[256];
  __shared__ LevelDev s_lev[R3D_MAX_LEVELS];
  __shared__ uint32_t s_hdr[BRICKS ? 64 : 1];  // bricks: the page header
  __shared__ uint32_t s_ts[R3D_MAX_VIEWS_PER_CALL + 1];
  __shared__ uint32_t s_tile;
  __shared__ uint4 s_red[NW][2];
  __shared__ uint32_t s_opq[8];  // TF intervals that can be visible (tf_visible)
  __shared__ float s_coop[SHADED_KIND<KIND> ? NW * COOP_FLOATS : 1];  // per wave: gradient taps
  const uint32_t lane = threadIdx.x;
  for (uint32_t i = lane; i < 256u; i += NT)
    s_tf[i] = make_float4(a.tf[i * 4 + 0], a.tf[i * 4 + 1], a.tf[i * 4 + 2], a.tf[i * 4 + 3]);
  if (!BRICKS && lane < a.vol.nlev) s_lev[lane] = a.vol.levels[lane];
  if (BRICKS)  // nlev held to 1..15: the per-level words stay inside the header
    for (uint32_t i = lane; i < 64u; i += NT) s_hdr[i] = i ? a.bk.page[i] : min(max(a.bk.page[0], 1u), 15u);
  for (uint32_t i = lane; i <= a.nviews; i += NT) s_ts[i] = a.tile_start[i];
  __syncthreads();
  if (lane < 64u) {  // wave 0: one bit per TF interval [i, i + 1] whose alpha is not zero at both ends
#pragma unroll
    for (uint32_t c = 0; c < 4u; c++) {
      const uint32_t i = c * 64u + lane, i1 = i < 255u ? i + 1u : 255u;
      const uint64_t m = __ballot(s_tf[i].w != 0.0f || s_tf[i1].w != 0.0f);
      if (lane == 0) {
        s_opq[2u * c] = (uint32_t)m;
        s_opq[2u * c + 1u] = (uint32_t)(m >> 32);
      }
    }
  }
  __syncthreads();
  const SamplerT<WIDE || MEASURE> S{a.vol.base, s_lev, a.vol.nlev, MEASURE ? a.touched : nullptr};
  uint32_t st = 0, sk = 0, sh = 0, nr = 0, bl = 0, bs = 0;
  uint32_t tile = blockIdx.x;
  for (;;) {
    if (a.queue) {
      if (lane == 0) s_tile = atomicAdd(a.queue, 1u);
      __syncthreads();
      tile = __builtin_amdgcn_readfirstlane(s_tile);
      __syncthreads();
    }
    if (tile >= a.ntiles) break;
    if (a.order) tile = a.order[tile];
    // tile -> view: the last v with tile_start[v] <= tile
    uint32_t lo = 0, hi = a.nviews;
    while (hi - lo > 1u) {
      const uint32_t mid = (lo + hi) >> 1;
      if (s_ts[mid] <= tile) lo = mid;
      else hi = mid;
    }
    const uint32_t v = __builtin_amdgcn_readfirstlane(lo);
    const r3d_frame_params &pc = a.views[v];
    const bool half = (pc.view_flags & 32u) != 0u;
    const uint32_t W = pc.viewport[0], H = pc.viewport[1];
    const uint32_t RW = half ? (W + 1u) / 2u : W, RH = half ? (H + 1u) / 2u : H;
    const uint32_t tiles_x = (RW + TW - 1u) / TW;
    const uint32_t lt = tile - s_ts[v];
    const uint32_t tx = (lt % tiles_x) * TW + (lane % TW);
    const uint32_t ty = (lt / tiles_x) * TH + (lane / TW);
    if (tx < RW && ty < RH) {
      const uint32_t lx = half ? tx * 2u : tx, ly = half ? ty * 2u : ty;
      float *coop = s_coop + (SHADED_KIND<KIND> ? (lane >> 6) * COOP_FLOATS : 0u);
      const RayOut R = BRICKS ? march_bricks<KIND>(a.bk, s_hdr, s_tf, s_opq, pc, a.cutoff, lx, ly, coop)
                              : march<KIND, MEASURE>(S, a.vol, s_tf, s_opq, pc, a.cutoff, lx, ly, coop);
      st += R.steps;
      sk += R.skips;
      sh += R.shaded;
      nr += 1u;
      bl += R.blends;
      bs += R.blends_shaded;
      const uint32_t ox = pc.view_org & 0xffffu, oy = pc.view_org >> 16;
      const uint32_t rep = half ? 2u : 1u;
      for (uint32_t dy = 0; dy < rep; dy++)
        for (uint32_t dx = 0; dx < rep; dx++) {
          const uint32_t qx = lx + dx, qy = ly + dy;
          if (qx >= W || qy >= H) continue;
          const uint32_t gx = ox + qx, gy = oy + qy;
          if (gx >= a.img_w || gy >= a.img_h) continue;
          const size_t pi = (size_t)gy * a.img_w + gx;
          if (a.img) ((float4 *)a.img)[pi] = make_float4(R.c[0], R.c[1], R.c[2], R.c[3]);
          if (a.img8) {
            uint32_t w = 0;
#pragma unroll
            for (int c = 0; c < 4; c++) w |= (uint32_t)(fsat(R.c[c]) * 255.0f + 0.5f) << (8 * c);
            ((uint32_t *)a.img8)[pi] = w;
          }
          if (a.px_steps) a.px_steps[pi] = R.steps;
        }
    }
    if (!a.queue) break;
  }
  if (a.partials) {  // two plain stores per workgroup; k_reduce_counters sums them
    st = wave_sum(st);
    sk = wave_sum(sk);
    sh = wave_sum(sh);
    nr = wave_sum(nr);
    bl = wave_sum(bl);
    bs = wave_sum(bs);
    if ((lane & 63u) == 0) {
      s_red[lane >> 6][0] = make_uint4(st, sk, sh, nr);
      s_red[lane >> 6][1] = make_uint4(bl, bs, 0u, 0u);
    }
    __syncthreads();
    if (lane < 2u) {
      uint4 t = s_red[0][lane];
#pragma unroll
      for (uint32_t w = 1; w < NW; w++) {
        t.x += s_red[w][lane].x;
        t.y += s_red[w][lane].y;
        t.z += s_red[w][lane].z;
        t.w += s_red[w][lane].w;
      }
      ((uint4 *)a.partials)[2u * blockIdx.x + lane] = t;
    }
  }
}

// An ordered 8x8 cube-mode launch of FULL, FAST, MIP or FLAT views, four
// waves per workgroup: ranks below tier2 (the costliest tiles of a shaded
// launch) render two tiles per workgroup with a lane pair per ray (each wave
// 32 rays, march_pair); the rest one tile per wave with one lane per ray.
// Four tiles share a workgroup's staged transfer function and level table.
// Pixels and counters are those of k_raycast.
template <int KIND, bool WIDE>
__global__ __launch_bounds__(256) void k_raycast_tiers(RayArgs a) {
  constexpr uint32_t NT = 256u, NW = 4u;
  __shared__ float4 s_tf[256];
  __shared__ LevelDev s_lev[R3D_MAX_LEVELS];
  __shared__ uint32_t s_ts[R3D_MAX_VIEWS_PER_CALL + 1];
  __shared__ uint4 s_red[NW][2];
  __shared__ uint32_t s_opq[8];
  __shared__ float s_coop[SHADED_KIND<KIND> ? NW * COOP_FLOATS : 1];
  const uint32_t lane = threadIdx.x, wv = lane >> 6, wl = lane & 63u;
  for (uint32_t i = lane; i < 256u; i += NT)
    s_tf[i] = make_float4(a.tf[i * 4 + 0], a.tf[i * 4 + 1], a.tf[i * 4 + 2], a.tf[i * 4 + 3]);
  if (lane < a.vol.nlev) s_lev[lane] = a.vol.levels[lane];
  for (uint32_t i = lane; i <= a.nviews; i += NT) s_ts[i] = a.tile_start[i];
  __syncthreads();
  if (lane < 64u) {  // the TF intervals whose alpha is not zero at both ends (as k_raycast)
#pragma unroll
    for (uint32_t c = 0; c < 4u; c++) {
      const uint32_t i = c * 64u + lane, i1 = i < 255u ? i + 1u : 255u;
      const uint64_t m = __ballot(s_tf[i].w != 0.0f || s_tf[i1].w != 0.0f);
      if (lane == 0) {
        s_opq[2u * c] = (uint32_t)m;
        s_opq[2u * c + 1u] = (uint32_t)(m >> 32);
      }
    }
  }
  __syncthreads();
  const SamplerT<WIDE> S{a.vol.base, s_lev, a.vol.nlev, nullptr};
  // this wave's tile rank, lanes per ray and first ray of the tile
  const uint32_t w = blockIdx.x, w2 = (a.tier2 + 1u) / 2u;
  uint32_t rank, lpr, ray0;
  if (w < w2) {
    rank = 2u * w + (wv >> 1);
    if (rank >= a.tier2) rank = a.ntiles;  // an odd tier: the last workgroup's second tile is not there
    lpr = 2u;
    ray0 = 32u * (wv & 1u);
  } else {
    rank = a.tier2 + 4u * (w - w2) + wv;
    lpr = 1u;
    ray0 = 0u;
  }
  uint32_t st = 0, sk = 0, sh = 0, nr = 0;
  if (rank < a.ntiles) {
    const uint32_t tile = a.order[rank];
    uint32_t lo = 0, hi = a.nviews;
    while (hi - lo > 1u) {
      const uint32_t mid = (lo + hi) >> 1;
      if (s_ts[mid] <= tile) lo = mid;
      else hi = mid;
    }
    const uint32_t v = __builtin_amdgcn_readfirstlane(lo);
    const r3d_frame_params &pc = a.views[v];
    const bool half = (pc.view_flags & 32u) != 0u;
    const uint32_t W = pc.viewport[0], H = pc.viewport[1];
    const uint32_t RW = half ? (W + 1u) / 2u : W, RH = half ? (H + 1u) / 2u : H;
    const uint32_t tiles_x = (RW + 7u) / 8u;
    const uint32_t lt = tile - s_ts[v];
    const uint32_t ray = ray0 + wl / lpr, j = wl % lpr;
    const uint32_t tx = (lt % tiles_x) * 8u + (ray & 7u);
    const uint32_t ty = (lt / tiles_x) * 8u + (ray >> 3);
    if (tx < RW && ty < RH) {
      const uint32_t lx = half ? tx * 2u : tx, ly = half ? ty * 2u : ty;
      float *coop = s_coop + (SHADED_KIND<KIND> ? wv * COOP_FLOATS : 0u);
      RayOut R;
      if constexpr (SHADED_KIND<KIND>) {
        R = lpr == 2u ? march_pair<KIND>(S, a.vol, s_tf, s_opq, pc, a.cutoff, lx, ly, j, coop)
                      : march<KIND, false>(S, a.vol, s_tf, s_opq, pc, a.cutoff, lx, ly, coop);
      } else {
        R = march<KIND, false>(S, a.vol, s_tf, s_opq, pc, a.cutoff, lx, ly, coop);
      }
      if (j == 0u) {
        st += R.steps;
        sk += R.skips;
        sh += R.shaded;
        nr += 1u;
        const uint32_t ox = pc.view_org & 0xffffu, oy = pc.view_org >> 16;
        const uint32_t rep = half ? 2u : 1u;
        for (uint32_t dy = 0; dy < rep; dy++)
          for (uint32_t dx = 0; dx < rep; dx++) {
            const uint32_t qx = lx + dx, qy = ly + dy;
            if (qx >= W || qy >= H) continue;
            const uint32_t gx = ox + qx, gy = oy + qy;
            if (gx >= a.img_w || gy >= a.img_h) continue;
            const size_t pi = (size_t)gy * a.img_w + gx;
            if (a.img) ((float4 *)a.img)[pi] = make_float4(R.c[0], R.c[1], R.c[2], R.c[3]);
            if (a.img8) {
              uint32_t w8 = 0;
#pragma unroll
              for (int c = 0; c < 4; c++) w8 |= (uint32_t)(fsat(R.c[c]) * 255.0f + 0.5f) << (8 * c);
              ((uint32_t *)a.img8)[pi] = w8;
            }
            if (a.px_steps) a.px_steps[pi] = R.steps;
          }
      }
    }
  }
  if (a.partials) {  // two plain stores per workgroup; k_reduce_counters sums them
    st = wave_sum(st);
    sk = wave_sum(sk);
    sh = wave_sum(sh);
    nr = wave_sum(nr);
    if (wl == 0u) {
      s_red[wv][0] = make_uint4(st, sk, sh, nr);
      s_red[wv][1] = make_uint4(0u, 0u, 0u, 0u);
    }
    __syncthreads();
    if (lane < 2u) {
      uint4 t = s_red[0][lane];
#pragma unroll
      for (uint32_t i = 1; i < NW; i++) {
        t.x += s_red[i][lane].x;
        t.y += s_red[i][lane].y;
        t.z += s_red[i][lane].z;
        t.w += s_red[i][lane].w;
      }
      ((uint4 *)a.partials)[2u * blockIdx.x + lane] = t;
    }
  }
}

// Sums `n` workgroup partials (two uint4 each) into counters[0..ncount):
// each workgroup reduces a strided share and adds it with 64-bit integer
// atomics (exact in any order).
__global__ __launch_bounds__(256) void k_reduce_counters(const uint32_t *partials, uint32_t n, uint32_t ncount,
                                                         unsigned long long *counters) {
  __shared__ unsigned long long red[6][256];
  unsigned long long acc[6] = {0, 0, 0, 0, 0, 0};
  for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < n; i += gridDim.x * 256u) {
    const uint4 p = ((const uint4 *)partials)[2u * i];
    const uint4 q = ((const uint4 *)partials)[2u * i + 1u];
    acc[0] += p.x;
    acc[1] += p.y;
    acc[2] += p.z;
    acc[3] += p.w;
    acc[4] += q.x;
    acc[5] += q.y;
  }
  for (int c = 0; c < 6; c++) red[c][threadIdx.x] = acc[c];
  __syncthreads();
  for (uint32_t s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s)
      for (int c = 0; c < 6; c++) red[c][threadIdx.x] += red[c][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x < ncount && red[threadIdx.x][0]) atomicAdd(&counters[threadIdx.x], red[threadIdx.x][0]);
}

static uint32_t reduce_grid(uint32_t n) {
  const uint32_t g = (n + 2047u) / 2048u;
  return g < 1u ? 1u : (g > 512u ? 512u : g);
}

__global__ void k_zero_u64(unsigned long long *p, uint32_t n) {
  if (threadIdx.x < n) p[threadIdx.x] = 0ull;
}

__global__ void k_zero_u32(uint32_t *p, uint32_t n) {
  if (threadIdx.x < n) p[threadIdx.x] = 0u;
}

// ---- launch order: the costliest tiles first ---------------------------
// Workgroups start in blockIdx order, and a frame ends when its slowest tile
// does. On the 1080p frame a tile's time follows its longest ray (up to ~490
// samples, ~0.6 ms for one tile), and in raster order the long central tiles
// start a third of the way into the frame, so the frame's second half runs
// a few of them on an almost idle chip. Sorting the launch by an estimate of
// each tile's march length starts them first; the short and empty tiles fill
// in behind. Every tile is still rendered exactly once and no pixel depends
// on the order.
constexpr uint32_t ORDER_BUCKETS = 256u;
// Tiles below this count (per launch) keep raster order: the two extra
// launches would cost more than the tail they remove.
constexpr uint32_t ORDER_MIN_TILES = 4096u;
// The costliest R3D_PAIR_TIER_NUM sixteenths of an ordered 8x8 FULL or FAST
// launch take a lane pair per ray (k_raycast_tiers), at most
// R3D_PAIR_TILES_PER_CU tiles per CU: the tiles that start first and bound
// the launch's tail; beyond them the pairs' extra work only costs time.
#ifndef R3D_PAIR_TIER_NUM
#define R3D_PAIR_TIER_NUM 4u
#endif
#ifndef R3D_PAIR_TILES_PER_CU
#define R3D_PAIR_TILES_PER_CU 8u
#endif
constexpr uint32_t ORDER_MAX_TILES = 1u << 24;  // a tile's rank inside its bucket takes 24 bits

__global__ void k_fill_u32(uint32_t *p, uint32_t n) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0u;
}

// Estimated samples of the ray through pixel (px, py): the box entry and exit
// of the march's own setup, then the step control's integral dt / (dt0 *
// lod_step(t)) by an 8-point midpoint rule with hardware log2/exp2 (an
// ordering key only; it shapes no pixel). ext: the volume box, md: its
// largest dimension in voxels, budget: the step budget of the span step.
__device__ float ray_cost(const r3d_frame_params &pc, float px, float py, V3 ext, float md, float max_lod,
                          float budget) {
  const V3 fwd = mk(pc.cam_forward[0], pc.cam_forward[1], pc.cam_forward[2]);
  const V3 right = mk(pc.cam_right[0], pc.cam_right[1], pc.cam_right[2]);
  const V3 up = mk(pc.cam_up[0], pc.cam_up[1], pc.cam_up[2]);
  const bool ortho = (pc.view_flags & 1u) != 0u;
  const float ndx = (px + 0.5f) / (float)pc.viewport[0] * 2.0f - 1.0f;
  const float ndy = 1.0f - (py + 0.5f) / (float)pc.viewport[1] * 2.0f;
  V3 dir = ortho ? vnorm(fwd) : vnorm(vadd(vadd(fwd, vscale(right, ndx)), vscale(up, ndy)));
  V3 org = mk(pc.cam_origin[0], pc.cam_origin[1], pc.cam_origin[2]);
  if (ortho) org = vadd(vadd(org, vscale(right, ndx)), vscale(up, ndy));
  const V3 ctr = vscale(ext, 0.5f);
  const V3 c0 = mk(pc.vol_r0[0], pc.vol_r1[0], pc.vol_r2[0]);
  const V3 c1 = mk(pc.vol_r0[1], pc.vol_r1[1], pc.vol_r2[1]);
  const V3 c2 = mk(pc.vol_r0[2], pc.vol_r1[2], pc.vol_r2[2]);
  const V3 q = vsub(vsub(org, mk(pc.vol_tx, pc.vol_ty, pc.vol_tz)), ctr);
  const V3 ro = vadd(mk(vdot(c0, q), vdot(c1, q), vdot(c2, q)), ctr);
  dir = mk(vdot(c0, dir), vdot(c1, dir), vdot(c2, dir));
  const V3 inv = mk(1.0f / dir.x, 1.0f / dir.y, 1.0f / dir.z);
  const V3 t0 = mk((0.0f - ro.x) * inv.x, (0.0f - ro.y) * inv.y, (0.0f - ro.z) * inv.z);
  const V3 t1 = mk((ext.x - ro.x) * inv.x, (ext.y - ro.y) * inv.y, (ext.z - ro.z) * inv.z);
  const float hx = fmaxf(fmaxf(fminf(t0.x, t1.x), fminf(t0.y, t1.y)), fminf(t0.z, t1.z));
  const float hy = fminf(fminf(fmaxf(t0.x, t1.x), fmaxf(t0.y, t1.y)), fmaxf(t0.z, t1.z));
  const float tn = fmaxf(hx, 0.0f), tfar = fminf(hy, 1e30f);
  if (!(tn < tfar)) return 0.0f;
  const float voxel = 1.0f / md, dt = fmaxf(pc.step_voxels, 1e-6f) * voxel;
  const float cone = 2.0f * vlen(up) / (float)pc.viewport[1];
  const float span_step = fmaxf((tfar - tn) / (budget * dt), 1.0f);
  const float h = (tfar - tn) * 0.125f;
  float n = 0.0f;
  for (int i = 0; i < 8; i++) {
    const float tt = tn + ((float)i + 0.5f) * h;
    const float fv = (ortho ? cone : tt * cone) * md;
    const float lod = fminf(fmaxf(__log2f(fmaxf(fv, 1.0f)) + pc.lod_bias, 0.0f), max_lod);
    n += h / (dt * fmaxf(exp2f(lod), span_step));
  }
  return n;
}

// A tile's bucket: the largest estimate over its corners and centre, log scale.
template <int TW, int TH, bool BRICKS>
__device__ uint32_t tile_bucket(const RayArgs &a, uint32_t tile) {
  uint32_t lo = 0, hi = a.nviews;
  while (hi - lo > 1u) {
    const uint32_t mid = (lo + hi) >> 1;
    if (a.tile_start[mid] <= tile) lo = mid;
    else hi = mid;
  }
  const r3d_frame_params &pc = a.views[lo];
  const bool half = (pc.view_flags & 32u) != 0u;
  const uint32_t W = pc.viewport[0], H = pc.viewport[1];
  const uint32_t RW = half ? (W + 1u) / 2u : W, RH = half ? (H + 1u) / 2u : H;
  const uint32_t tiles_x = (RW + TW - 1u) / TW;
  const uint32_t lt = tile - a.tile_start[lo];
  const uint32_t x0 = (lt % tiles_x) * TW, y0 = (lt / tiles_x) * TH;
  const uint32_t x1 = min(x0 + TW - 1u, RW - 1u), y1 = min(y0 + TH - 1u, RH - 1u);
  V3 ext = mk(1.0f, 1.0f, 1.0f);
  float md, max_lod, budget;
  if (BRICKS) {
    const uint32_t nl = min(max(a.bk.page[0], 1u), 15u);
    const uint32_t sx = a.bk.page[1], sy = a.bk.page[2], sz = a.bk.page[3];
    const uint32_t m = max(max(sx, sy), max(sz, 1u));
    md = (float)m;
    ext = mk((float)sx / md, (float)sy / md, (float)sz / md);
    max_lod = (float)(nl - 1u);
    budget = pc.mode == MODE_FULL ? 192.0f : 256.0f;
  } else {
    md = (float)max(a.vol.maxdim, 1u);
    max_lod = fminf(pc.max_mip, (float)(a.vol.nlev > 0u ? a.vol.nlev - 1u : 0u));
    budget = (float)MAX_STEPS;
  }
  const float sc = half ? 2.0f : 1.0f;
  const float xs[5] = {(float)x0, (float)x1, (float)x0, (float)x1, (float)((x0 + x1) / 2u)};
  const float ys[5] = {(float)y0, (float)y0, (float)y1, (float)y1, (float)((y0 + y1) / 2u)};
  float c = 0.0f;
#pragma unroll
  for (int i = 0; i < 5; i++) c = fmaxf(c, ray_cost(pc, xs[i] * sc, ys[i] * sc, ext, md, max_lod, budget));
  return c > 0.0f ? min((uint32_t)(__log2f(1.0f + c) * 24.0f), ORDER_BUCKETS - 1u) : 0u;
}

// Pass 1 of the order: each tile's key (the largest estimate over its
// corners and centre, on a log scale in ORDER_BUCKETS buckets) and its rank
// inside its bucket. code[tile] = bucket << 24 | rank. Ranks come from an
// LDS histogram per workgroup plus one global atomic per occupied bucket (a
// global atomic per tile serialises: most tiles of a frame share the empty
// bucket).
template <int TW, int TH, bool BRICKS>
__global__ __launch_bounds__(256) void k_tile_cost(RayArgs a, uint32_t *code, uint32_t *counts) {
  __shared__ uint32_t hist[ORDER_BUCKETS], base[ORDER_BUCKETS];
  const uint32_t tile = blockIdx.x * 256u + threadIdx.x;
  hist[threadIdx.x] = 0u;
  __syncthreads();
  uint32_t b = 0u, lr = 0u;
  if (tile < a.ntiles) {
    b = tile_bucket<TW, TH, BRICKS>(a, tile);
    lr = atomicAdd(&hist[b], 1u);
  }
  __syncthreads();
  const uint32_t hc = hist[threadIdx.x];
  if (hc) base[threadIdx.x] = atomicAdd(&counts[threadIdx.x], hc);
  __syncthreads();
  if (tile < a.ntiles) code[tile] = b << 24 | (base[b] + lr);
}

// Pass 2: order[off(bucket) + rank] = tile, the buckets from the costliest
// down (each workgroup rebuilds the 256 bucket offsets in LDS).
__global__ __launch_bounds__(256) void k_tile_scatter(const uint32_t *code, const uint32_t *counts,
                                                      uint32_t ntiles, uint32_t *order) {
  __shared__ uint32_t sc[ORDER_BUCKETS];
  const uint32_t t = threadIdx.x;
  const uint32_t v = counts[ORDER_BUCKETS - 1u - t];  // descending bucket order
  sc[t] = v;
  __syncthreads();
  for (uint32_t d = 1; d < ORDER_BUCKETS; d <<= 1) {
    const uint32_t x = t >= d ? sc[t - d] : 0u;
    __syncthreads();
    sc[t] += x;
    __syncthreads();
  }
  const uint32_t excl = sc[t] - v;  // tiles in costlier buckets than bucket 255 - t
  __syncthreads();
  sc[ORDER_BUCKETS - 1u - t] = excl;
  __syncthreads();
  const uint32_t tile = blockIdx.x * 256u + t;
  if (tile < ntiles) {
    const uint32_t c = code[tile];
    order[sc[c >> 24] + (c & 0xffffffu)] = tile;
  }
}

typedef void (*RayKernel)(RayArgs);
#define R3D_RAY_SHAPES(K, B, W)                                                                         \
  {                                                                                                     \
    k_raycast<K, 8, 8, B, false, W>, k_raycast<K, 16, 4, B, false, W>, k_raycast<K, 16, 8, B, false, W>, \
        k_raycast<K, 16, 16, B, false, W>                                                               \
  }
// The 8x8 cube-mode measurement instances (blend counts, touched bricks).
static const RayKernel measure_kernels[NKINDS] = {k_raycast<KIND_FULL, 8, 8, false, true>,
                                                k_raycast<KIND_MIP, 8, 8, false, true>,
                                                k_raycast<KIND_FLAT, 8, 8, false, true>,
                                                k_raycast<KIND_OTHER, 8, 8, false, true>,
                                                k_raycast<KIND_FAST, 8, 8, false, true>};
// Rows: cube mode with 32-bit offsets (a pyramid below 4 GiB), bricks-LOD,
// cube mode with 64-bit offsets.
static const RayKernel ray_kernels[3][NKINDS][NSHAPES] = {
    {R3D_RAY_SHAPES(KIND_FULL, false, false), R3D_RAY_SHAPES(KIND_MIP, false, false),
     R3D_RAY_SHAPES(KIND_FLAT, false, false), R3D_RAY_SHAPES(KIND_OTHER, false, false),
     R3D_RAY_SHAPES(KIND_FAST, false, false)},
    {R3D_RAY_SHAPES(KIND_FULL, true, true), R3D_RAY_SHAPES(KIND_MIP, true, true),
     R3D_RAY_SHAPES(KIND_FLAT, true, true), R3D_RAY_SHAPES(KIND_OTHER, true, true),
     R3D_RAY_SHAPES(KIND_FAST, true, true)},
    {R3D_RAY_SHAPES(KIND_FULL, false, true), R3D_RAY_SHAPES(KIND_MIP, false, true),
     R3D_RAY_SHAPES(KIND_FLAT, false, true), R3D_RAY_SHAPES(KIND_OTHER, false, true),
     R3D_RAY_SHAPES(KIND_FAST, false, true)}};
#undef R3D_RAY_SHAPES
// k_raycast_tiers by offset width (32-bit, 64-bit) and kind; none for the
// remaining modes, which take k_raycast.
static const RayKernel tier_kernels[2][NKINDS] = {
    {k_raycast_tiers<KIND_FULL, false>, k_raycast_tiers<KIND_MIP, false>, k_raycast_tiers<KIND_FLAT, false>,
     nullptr, k_raycast_tiers<KIND_FAST, false>},
    {k_raycast_tiers<KIND_FULL, true>, k_raycast_tiers<KIND_MIP, true>, k_raycast_tiers<KIND_FLAT, true>, nullptr,
     k_raycast_tiers<KIND_FAST, true>}};
typedef void (*CostKernel)(RayArgs, uint32_t *, uint32_t *);
static const CostKernel cost_kernels[2][NSHAPES] = {
    {k_tile_cost<8, 8, false>, k_tile_cost<16, 4, false>, k_tile_cost<16, 8, false>, k_tile_cost<16, 16, false>},
    {k_tile_cost<8, 8, true>, k_tile_cost<16, 4, true>, k_tile_cost<16, 8, true>, k_tile_cost<16, 16, true>}};

static int kind_of(uint32_t mode, uint32_t quality) {
  switch (mode) {
  case MODE_FULL: return quality == R3D_QUALITY_FAST ? KIND_FAST : KIND_FULL;
  case MODE_MIP: return KIND_MIP;
  case MODE_FLAT: return KIND_FLAT;
  default: return KIND_OTHER;
  }
}

// Views grouped by kernel kind, with one tile prefix table per shape.
// Device layout: [views by kind][NSHAPES x (nviews + NKINDS) tile starts].
struct ViewGroups {
  uint32_t nviews = 0;
  uint32_t gstart[NKINDS + 1] = {};
  uint32_t tpos[NKINDS] = {};             // offset of kind k's table inside a shape's table
  uint32_t gtiles[NSHAPES][NKINDS] = {};  // tiles of kind k at shape s
  size_t view_bytes() const { return sizeof(r3d_frame_params) * nviews; }
  size_t table_words() const { return (size_t)nviews + NKINDS; }
  size_t bytes() const { return view_bytes() + sizeof(uint32_t) * table_words() * NSHAPES; }
};

static void group_views(const r3d_frame_params *views, uint32_t nviews, uint32_t quality, uint8_t *buf,
                        ViewGroups *g) {
  g->nviews = nviews;
  r3d_frame_params *pv = (r3d_frame_params *)buf;
  uint32_t *tab = (uint32_t *)(buf + g->view_bytes());
  uint32_t n = 0, tcur = 0;
  uint32_t tiles[NSHAPES];
  for (int k = 0; k < NKINDS; k++) {
    g->gstart[k] = n;
    g->tpos[k] = tcur;
    for (int s = 0; s < NSHAPES; s++) tiles[s] = 0;
    for (uint32_t i = 0; i < nviews; i++) {
      if (kind_of(views[i].mode, quality) != k) continue;
      pv[n] = views[i];
      const bool half = (views[i].view_flags & 32u) != 0u;
      const uint32_t RW = half ? (views[i].viewport[0] + 1u) / 2u : views[i].viewport[0];
      const uint32_t RH = half ? (views[i].viewport[1] + 1u) / 2u : views[i].viewport[1];
      for (int s = 0; s < NSHAPES; s++) {
        tab[s * g->table_words() + tcur] = tiles[s];
        tiles[s] += ((RW + SHAPE_W[s] - 1u) / SHAPE_W[s]) * ((RH + SHAPE_H[s] - 1u) / SHAPE_H[s]);
      }
      tcur++;
      n++;
    }
    for (int s = 0; s < NSHAPES; s++) {
      tab[s * g->table_words() + tcur] = tiles[s];
      g->gtiles[s][k] = tiles[s];
    }
    tcur++;
  }
  g->gstart[NKINDS] = n;
}

// Counter partial slots one launch of `g` at tuning `t` writes.
static uint32_t partial_slots(const r3d_hip_ctx *ctx, const ViewGroups &g, const r3d_raycast_tuning &t) {
  const int s = shape_index(t.tile_w, t.tile_h);
  uint32_t n = 0;
  for (int k = 0; k < NKINDS; k++) {
    const uint32_t tiles = g.gtiles[s][k];
    const uint32_t cap = (uint32_t)(ctx->num_cus > 0 ? ctx->num_cus : 256) * t.blocks_per_cu;
    n += t.blocks_per_cu && cap < tiles ? cap : tiles;
  }
  return n;
}

static int check_tuning(r3d_hip_ctx *ctx, const r3d_raycast_tuning *t) {
  if (!t) return r3d_fail(ctx, R3D_ERR_INVALID, "raycast tuning is NULL");
  if (shape_index(t->tile_w, t->tile_h) < 0)
    return r3d_fail(ctx, R3D_ERR_INVALID, "raycast tuning: tile %ux%u is not 8x8, 16x4, 16x8 or 16x16",
                    t->tile_w, t->tile_h);
  if (t->blocks_per_cu > MAX_BLOCKS_PER_CU)
    return r3d_fail(ctx, R3D_ERR_INVALID, "raycast tuning: blocks_per_cu %u above %u", t->blocks_per_cu,
                    MAX_BLOCKS_PER_CU);
  return R3D_OK;
}

// Words of launch-order scratch a launch of `g` at any tuning needs: the
// bucket counters of every kind, then each tile's code and its order entry.
static size_t order_words(const ViewGroups &g) {
  uint32_t most = 0;
  for (int s = 0; s < NSHAPES; s++) {
    uint32_t n = 0;
    for (int k = 0; k < NKINDS; k++) n += g.gtiles[s][k];
    most = n > most ? n : most;
  }
  return (size_t)NKINDS * ORDER_BUCKETS + 2u * (size_t)most;
}

// Kernel launches only (capture-safe): queue reset, one launch per kind,
// counter reduction into ncount counters. `partials` holds partial_slots()
// entries of two uint4. ncount 6 runs the 8x8 measurement instances (t must
// then be 8x8, one workgroup per tile). `order` (order_words(g) words, or
// NULL for raster order) holds the launch order of kinds with at least
// ORDER_MIN_TILES tiles, built by two small kernels ahead of each launch.
static int launch_groups(r3d_hip_ctx *ctx, RayArgs a, const r3d_frame_params *dviews, const uint32_t *dtables,
                         const ViewGroups &g, const r3d_raycast_tuning &t, uint32_t *partials, uint32_t *queue,
                         uint32_t *order, unsigned long long *counters, uint32_t ncount, hipStream_t s,
                         bool bricks = false) {
  const bool measure = ncount == 6u;
  const int sh = shape_index(t.tile_w, t.tile_h);
  const uint32_t nt = SHAPE_W[sh] * SHAPE_H[sh];
  if (t.blocks_per_cu) hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, s, queue, (uint32_t)NKINDS);
  uint32_t total = 0;
  bool ordered = false;
  for (int k = 0; k < NKINDS; k++) {
    total += g.gtiles[sh][k];
    ordered = ordered || (order && g.gstart[k + 1] > g.gstart[k] && g.gtiles[sh][k] >= ORDER_MIN_TILES &&
                          g.gtiles[sh][k] <= ORDER_MAX_TILES);
  }
  uint32_t *codes = order ? order + NKINDS * ORDER_BUCKETS : nullptr;
  uint32_t *perm = order ? codes + total : nullptr;
  if (ordered) {
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(256), 0, s, order, NKINDS * ORDER_BUCKETS);
    R3D_HIP_CHECK(ctx, hipGetLastError());
  }
  uint32_t pbase = 0, tbase = 0;
  for (int k = 0; k < NKINDS; k++) {
    const uint32_t nv = g.gstart[k + 1] - g.gstart[k];
    const uint32_t tiles = g.gtiles[sh][k];
    const uint32_t tb = tbase;
    tbase += tiles;
    if (nv == 0 || tiles == 0) continue;
    const bool ordered_k = order && tiles >= ORDER_MIN_TILES && tiles <= ORDER_MAX_TILES;
    const RayKernel tier = ordered_k && !measure && !bricks && !t.blocks_per_cu && sh == 0
                               ? tier_kernels[a.vol.narrow ? 0 : 1][k]
                               : nullptr;
    uint32_t grid = tiles, threads = nt;
    if (tier) {  // shaded launches: the costliest ranks with a lane pair per ray (k_raycast_tiers)
      uint32_t pairs = 0;
      if (k == KIND_FULL || k == KIND_FAST) {
        const uint64_t frac = ((uint64_t)tiles * R3D_PAIR_TIER_NUM + 15u) / 16u;
        const uint64_t cap = (uint64_t)(ctx->num_cus > 0 ? ctx->num_cus : 256) * R3D_PAIR_TILES_PER_CU;
        pairs = (uint32_t)(frac < cap ? frac : cap);
      }
      a.tier2 = pairs;
      grid = (a.tier2 + 1u) / 2u + (tiles - a.tier2 + 3u) / 4u;
      threads = 256u;
    }
    if (t.blocks_per_cu) {
      const uint32_t cap = (uint32_t)(ctx->num_cus > 0 ? ctx->num_cus : 256) * t.blocks_per_cu;
      if (cap < grid) grid = cap;
    }
    a.partials = counters ? partials + (size_t)pbase * 8u : nullptr;
    pbase += grid;
    a.views = dviews + g.gstart[k];
    a.tile_start = dtables + (size_t)sh * g.table_words() + g.tpos[k];
    a.nviews = nv;
    a.ntiles = tiles;
    a.queue = t.blocks_per_cu ? queue + k : nullptr;
    a.order = nullptr;
    if (ordered_k) {
      uint32_t *ck = order + (size_t)k * ORDER_BUCKETS;
      const uint32_t og = (tiles + 255u) / 256u;
      hipLaunchKernelGGL(cost_kernels[bricks ? 1 : 0][sh], dim3(og), dim3(256), 0, s, a, codes + tb, ck);
      hipLaunchKernelGGL(k_tile_scatter, dim3(og), dim3(256), 0, s, (const uint32_t *)(codes + tb),
                         (const uint32_t *)ck, tiles, perm + tb);
      R3D_HIP_CHECK(ctx, hipGetLastError());
      a.order = perm + tb;
    }
    const int row = bricks ? 1 : a.vol.narrow ? 0 : 2;
    hipLaunchKernelGGL(measure ? measure_kernels[k] : tier ? tier : ray_kernels[row][k][sh], dim3(grid),
                       dim3(threads), 0, s, a);
    R3D_HIP_CHECK(ctx, hipGetLastError());
  }
  if (counters && pbase) {
    hipLaunchKernelGGL(k_reduce_counters, dim3(reduce_grid(pbase)), dim3(256), 0, s, (const uint32_t *)partials,
                       pbase, ncount, counters);
    R3D_HIP_CHECK(ctx, hipGetLastError());
  }
  return R3D_OK;
}

// abpa = 0: cube-mode views; otherwise bricks-LOD views of an atlas with abpa
// slots per axis (brick_mode = abpa | BR_LOD_FLAG, raycast.slang:157-161).
static int validate_views(r3d_hip_ctx *ctx, const r3d_frame_params *views, uint32_t nviews,
                          const r3d_hip_image *out, uint32_t abpa = 0) {
  if (!views || nviews == 0 || nviews > R3D_MAX_VIEWS_PER_CALL)
    return r3d_fail(ctx, R3D_ERR_INVALID, "raycast: nviews %u outside 1..%u", nviews, R3D_MAX_VIEWS_PER_CALL);
  if (!out || (!out->rgba && !out->rgba8) || out->width == 0 || out->height == 0)
    return r3d_fail(ctx, R3D_ERR_INVALID, "raycast: no output image");
  if (out->width > 65535u || out->height > 65535u)
    return r3d_fail(ctx, R3D_ERR_RANGE, "raycast: image above 65535 pixels per side");
  for (uint32_t i = 0; i < nviews; i++) {
    const r3d_frame_params &p = views[i];
    if (abpa) {
      if (!(p.brick_mode & 0x20000u) || (p.brick_mode & 0xffu) != abpa || p.slab_grid || p.clip_valid ||
          (p.view_flags & 2u))
        return r3d_fail(ctx, R3D_ERR_UNSUPPORTED, "raycast: view %u is not a bricks-LOD view of this atlas", i);
    } else if (p.slab_grid || p.clip_valid || p.brick_mode || (p.view_flags & 2u)) {
      return r3d_fail(ctx, R3D_ERR_UNSUPPORTED, "raycast: view %u is not a cube-mode view", i);
    }
    if (p.mode > MODE_FLAT) return r3d_fail(ctx, R3D_ERR_INVALID, "raycast: view %u mode %u", i, p.mode);
    if (p.viewport[0] == 0 || p.viewport[1] == 0 || p.viewport[0] > 65535u || p.viewport[1] > 65535u)
      return r3d_fail(ctx, R3D_ERR_INVALID, "raycast: view %u viewport %ux%u", i, p.viewport[0], p.viewport[1]);
    if (!(p.step_voxels > 0.0f)) return r3d_fail(ctx, R3D_ERR_INVALID, "raycast: view %u step_voxels", i);
    if (!abpa && !(p.max_mip >= 0.0f)) return r3d_fail(ctx, R3D_ERR_INVALID, "raycast: view %u max_mip", i);
  }
  return R3D_OK;
}

static RayArgs base_args(r3d_hip_ctx *ctx, const r3d_hip_volume *vol, const r3d_hip_image *out,
                         const r3d_raycast_opts *opts) {
  RayArgs a{};
  if (vol) a.vol = vol_dev(vol);  // a bricks-LOD render has no cube volume (nlev 0)
  a.tf = opts && opts->tf ? opts->tf : ctx->d_tf;
  a.img = out->rgba;
  a.img8 = out->rgba8;
  a.img_w = out->width;
  a.img_h = out->height;
  a.cutoff = opts && opts->alpha_cutoff > 0.0f ? opts->alpha_cutoff : 0.98f;
  a.px_steps = opts ? opts->px_steps : nullptr;
  a.touched = opts ? opts->touched : nullptr;
  return a;
}

// Copies `n` bytes into the next staging slot (pinned host + device buffer);
// the caller records the slot's event after the launches that read it.
int stage(r3d_hip_ctx *ctx, const void *src, size_t n, hipStream_t s, ViewSlot **slot_out) {
  ViewSlot &sl = ctx->slots[ctx->slot_next];
  ctx->slot_next = (ctx->slot_next + 1) % 4;
  if (sl.recorded) {
    R3D_HIP_CHECK(ctx, hipEventSynchronize(sl.ev));
    sl.recorded = false;
  }
  if (!sl.ev) R3D_HIP_CHECK(ctx, hipEventCreateWithFlags(&sl.ev, hipEventDisableTiming));
  if (sl.cap < n) {
    if (sl.host) hipHostFree(sl.host);
    if (sl.dev) hipFree(sl.dev);
    sl.host = sl.dev = nullptr;
    sl.cap = 0;
    size_t cap = n < 4096 ? 4096 : n;
    if (hipHostMalloc(&sl.host, cap, 0) != hipSuccess || hipMalloc(&sl.dev, cap) != hipSuccess)
      return r3d_fail(ctx, R3D_ERR_NOMEM, "raycast: staging %zu bytes", cap);
    sl.cap = cap;
  }
  memcpy(sl.host, src, n);
  R3D_HIP_CHECK(ctx, hipMemcpyAsync(sl.dev, sl.host, n, hipMemcpyHostToDevice, s));
  *slot_out = &sl;
  return R3D_OK;
}

}  // namespace r3d

using namespace r3d;

struct r3d_hip_raycast_plan {
  r3d_hip_ctx *ctx = nullptr;
  const r3d_hip_volume *vol = nullptr;
  ViewGroups g;
  void *d_mem = nullptr;  // views, tile tables, queue counters, counter partials
  r3d_frame_params *d_views = nullptr;
  uint32_t *d_tables = nullptr;
  uint32_t *d_queue = nullptr;
  uint32_t *d_partials = nullptr;
  uint32_t *d_order = nullptr;
  r3d_hip_image out{};
  r3d_raycast_opts opts{};
  r3d_raycast_tuning tuning{8u, 8u, 0u, 0.0f};
};

// r3d_hip_raycast and r3d_hip_bricks_raycast: one of vol / bk is given.
// FNV-1a over bytes, continuing from k.
static uint64_t fnv1a(uint64_t k, const void *p, size_t n) {
  const uint8_t *b = (const uint8_t *)p;
  for (size_t i = 0; i < n; i++) k = (k ^ b[i]) * 1099511628211ull;
  return k;
}

// The pane cache (vkbackend.c:7066-7079, 7144-7159): the key of view v is
// every parameter byte but frame_index (a cached view keeps its last jitter),
// the scene generation, the handle, the outputs, the options and the layout
// of all views (a layout change redraws every view). Returns how many views
// changed; their indices go to draw[], their keys to keys[].
static uint32_t pane_select(const r3d_hip_ctx *ctx, const void *handle, const r3d_frame_params *views,
                            uint32_t nviews, const r3d_hip_image *out, const r3d_raycast_opts *opts,
                            uint32_t *draw, uint64_t *keys) {
  uint64_t lay = 1469598103934665603ull ^ nviews;
  for (uint32_t v = 0; v < nviews; v++) {
    lay = fnv1a(lay, &views[v].view_org, sizeof views[v].view_org);
    lay = fnv1a(lay, views[v].viewport, sizeof views[v].viewport);
  }
  struct {
    uint64_t gen, lay;
    const void *handle, *rgba, *rgba8, *tf, *px_steps, *touched;
    uint32_t width, height, flags;
    float alpha_cutoff;
  } common{ctx->scene_gen, lay, handle, out->rgba, out->rgba8, opts ? opts->tf : nullptr,
           opts ? opts->px_steps : nullptr, opts ? opts->touched : nullptr, out->width, out->height,
           opts ? opts->flags : 0u, opts ? opts->alpha_cutoff : 0.98f};
  const uint64_t k0 = fnv1a(1469598103934665603ull, &common, sizeof common);
  uint32_t n = 0;
  for (uint32_t v = 0; v < nviews; v++) {
    r3d_frame_params hp = views[v];
    hp.frame_index = 0;
    uint64_t k = fnv1a(k0, &hp, sizeof hp);
    k = (k ^ v) * 1099511628211ull;
    if (k == 0) k = 1;
    if (ctx->pane_key[v] == k) continue;  // unchanged: the output keeps its pixels
    draw[n] = v;
    keys[n++] = k;
  }
  return n;
}

static int raycast_impl(r3d_hip_ctx *ctx, const r3d_hip_volume *vol, const r3d_hip_bricks *bk,
                        const r3d_frame_params *views, uint32_t nviews, const r3d_hip_image *out,
                        const r3d_raycast_opts *opts, r3d_hip_stats *stats, void *stream) {
  int rc = validate_views(ctx, views, nviews, out, bk ? bk->desc.abpa : 0u);
  if (rc != R3D_OK) return rc;
  R3D_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  hipStream_t s = (hipStream_t)stream;
  if (stats) *stats = r3d_hip_stats{};
  // the pane cache: only views whose inputs changed are drawn
  const bool cache = opts && (opts->flags & R3D_RAYCAST_PANE_CACHE);
  uint32_t draw[R3D_MAX_VIEWS_PER_CALL];
  uint64_t keys[R3D_MAX_VIEWS_PER_CALL];
  std::vector<r3d_frame_params> sel;
  if (cache) {
    const uint32_t n = pane_select(ctx, bk ? (const void *)bk : (const void *)vol, views, nviews, out, opts, draw,
                                   keys);
    if (n == 0) return R3D_OK;
    if (n < nviews) {
      sel.resize(n);
      for (uint32_t i = 0; i < n; i++) sel[i] = views[draw[i]];
      views = sel.data();
      nviews = n;
    }
  } else {
    r3d_panes_reset(ctx);  // this call may overwrite what cached views left
  }
  uint8_t buf[sizeof(r3d_frame_params) * R3D_MAX_VIEWS_PER_CALL +
              sizeof(uint32_t) * (R3D_MAX_VIEWS_PER_CALL + NKINDS) * NSHAPES];
  ViewGroups g;
  group_views(views, nviews, ctx->quality, buf, &g);
  ViewSlot *slot = nullptr;
  rc = stage(ctx, buf, g.bytes(), s, &slot);
  if (rc != R3D_OK) return rc;
  unsigned long long *counters = opts ? (unsigned long long *)opts->counters : nullptr;
  const bool measure = opts && (opts->flags & R3D_RAYCAST_MEASURE);
  const uint32_t ncount = measure ? 6u : 4u;
  if (stats) {
    counters = ctx->d_counters;
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, s, counters, 6u);
  }
  // measuring: the 8x8 measurement instances, one workgroup per tile
  const r3d_raycast_tuning tuning = measure ? r3d_raycast_tuning{8u, 8u, 0u, 0.0f} : ctx->tuning;
  const uint32_t slots = partial_slots(ctx, g, tuning);
  if (counters && ctx->partials_cap < slots) {
    if (ctx->d_partials) {
      R3D_HIP_CHECK(ctx, hipStreamSynchronize(s));
      R3D_HIP_CHECK(ctx, hipFree(ctx->d_partials));
      ctx->d_partials = nullptr;
      ctx->partials_cap = 0;
    }
    const size_t cap = slots < 65536u ? 65536u : slots;
    if (hipMalloc(&ctx->d_partials, cap * 32u) != hipSuccess)
      return r3d_fail(ctx, R3D_ERR_NOMEM, "raycast: counter partials for %zu workgroups", cap);
    ctx->partials_cap = cap;
  }
  const size_t owords = order_words(g);
  if (ctx->order_cap < owords) {
    if (ctx->d_order) {
      R3D_HIP_CHECK(ctx, hipStreamSynchronize(s));
      uint32_t *old = ctx->d_order;
      ctx->d_order = nullptr;
      ctx->order_cap = 0;
      R3D_HIP_CHECK(ctx, hipFree(old));
    }
    const size_t cap = owords < (size_t)1 << 20 ? (size_t)1 << 20 : owords;
    if (hipMalloc(&ctx->d_order, cap * 4u) != hipSuccess)
      return r3d_fail(ctx, R3D_ERR_NOMEM, "raycast: launch order for %zu words", cap);
    ctx->order_cap = cap;
  }
  const r3d_frame_params *dviews = (const r3d_frame_params *)slot->dev;
  const uint32_t *dtables = (const uint32_t *)((const uint8_t *)slot->dev + g.view_bytes());
  RayArgs a = base_args(ctx, vol, out, opts);
  if (bk) a.bk = BricksDev{bk->d_page, bk->d_atlas, bk->desc.abpa, bk->page_words, bk->nslots};
  if (measure && bk) return r3d_fail(ctx, R3D_ERR_UNSUPPORTED, "bricks_raycast: no measurement mode");
  if (a.touched && !measure)
    return r3d_fail(ctx, R3D_ERR_INVALID, "raycast: a touched-brick bitmap needs R3D_RAYCAST_MEASURE");
  if (stats) R3D_HIP_CHECK(ctx, hipEventRecord(ctx->ev_ray[0], s));
  rc = launch_groups(ctx, a, dviews, dtables, g, tuning, ctx->d_partials, ctx->d_queue, ctx->d_order, counters,
                     ncount, s,
                     bk != nullptr);
  if (rc != R3D_OK) return rc;
  if (stats) R3D_HIP_CHECK(ctx, hipEventRecord(ctx->ev_ray[1], s));
  R3D_HIP_CHECK(ctx, hipEventRecord(slot->ev, s));
  slot->recorded = true;
  if (cache)
    for (uint32_t i = 0; i < nviews; i++) ctx->pane_key[draw[i]] = keys[i];
  if (stats) {
    unsigned long long h[6];
    R3D_HIP_CHECK(ctx, hipMemcpyAsync(h, counters, sizeof h, hipMemcpyDeviceToHost, s));
    R3D_HIP_CHECK(ctx, hipStreamSynchronize(s));
    float ms = 0.0f;
    R3D_HIP_CHECK(ctx, hipEventElapsedTime(&ms, ctx->ev_ray[0], ctx->ev_ray[1]));
    stats->samples = h[0];
    stats->skips = h[1];
    stats->shaded = h[2];
    stats->rays = h[3];
    stats->blended = h[4];
    stats->blended_shaded = h[5];
    stats->gpu_raycast_ns = (uint64_t)((double)ms * 1e6 + 0.5);
    stats->panes_drawn = nviews;
  }
  return R3D_OK;
}

extern "C" int r3d_hip_raycast(r3d_hip_ctx *ctx, const r3d_hip_volume *vol, const r3d_frame_params *views,
                               uint32_t nviews, const r3d_hip_image *out, const r3d_raycast_opts *opts,
                               r3d_hip_stats *stats, void *stream) {
  if (!ctx) return R3D_ERR_INVALID;
  if (!vol || vol->ctx != ctx) return r3d_fail(ctx, R3D_ERR_INVALID, "raycast: bad volume");
  return raycast_impl(ctx, vol, nullptr, views, nviews, out, opts, stats, stream);
}

extern "C" int r3d_hip_bricks_raycast(r3d_hip_ctx *ctx, const r3d_hip_bricks *b, const r3d_frame_params *views,
                                      uint32_t nviews, const r3d_hip_image *out, const r3d_raycast_opts *opts,
                                      r3d_hip_stats *stats, void *stream) {
  if (!ctx) return R3D_ERR_INVALID;
  if (!b || b->ctx != ctx) return r3d_fail(ctx, R3D_ERR_INVALID, "bricks_raycast: bad bricks handle");
  return raycast_impl(ctx, nullptr, b, views, nviews, out, opts, stats, stream);
}

// r3d_set_quality (src/render/render.h:241-242; src/vk/vkbackend.c:2075-2077):
// an unknown quality selects FULL.
extern "C" void r3d_hip_set_quality(r3d_hip_ctx *ctx, uint32_t quality) {
  if (!ctx) return;
  ctx->quality = quality == R3D_QUALITY_FAST ? R3D_QUALITY_FAST : R3D_QUALITY_FULL;
  r3d_scene_changed(ctx);
}

extern "C" int r3d_hip_set_raycast_tuning(r3d_hip_ctx *ctx, const r3d_raycast_tuning *t) {
  if (!ctx) return R3D_ERR_INVALID;
  const int rc = check_tuning(ctx, t);
  if (rc != R3D_OK) return rc;
  ctx->tuning = *t;
  r3d_scene_changed(ctx);
  return R3D_OK;
}

extern "C" int r3d_hip_raycast_plan_create(r3d_hip_ctx *ctx, const r3d_hip_volume *vol,
                                           const r3d_frame_params *views, uint32_t nviews,
                                           const r3d_hip_image *out, const r3d_raycast_opts *opts,
                                           r3d_hip_raycast_plan **plan) {
  if (!ctx) return R3D_ERR_INVALID;
  if (!plan) return r3d_fail(ctx, R3D_ERR_INVALID, "raycast plan: NULL output");
  *plan = nullptr;
  if (!vol || vol->ctx != ctx) return r3d_fail(ctx, R3D_ERR_INVALID, "raycast plan: bad volume");
  if (opts && (opts->touched || (opts->flags & R3D_RAYCAST_MEASURE)))
    return r3d_fail(ctx, R3D_ERR_UNSUPPORTED, "raycast plan: the measurement mode is for r3d_hip_raycast");
  int rc = validate_views(ctx, views, nviews, out);
  if (rc != R3D_OK) return rc;
  R3D_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  r3d_hip_raycast_plan *p = new (std::nothrow) r3d_hip_raycast_plan;
  if (!p) return r3d_fail(ctx, R3D_ERR_NOMEM, "raycast plan");
  p->ctx = ctx;
  p->vol = vol;
  p->out = *out;
  if (opts) p->opts = *opts;
  p->tuning = ctx->tuning;
  uint8_t *buf = (uint8_t *)malloc(sizeof(r3d_frame_params) * nviews +
                                   sizeof(uint32_t) * ((size_t)nviews + NKINDS) * NSHAPES);
  if (!buf) {
    delete p;
    return r3d_fail(ctx, R3D_ERR_NOMEM, "raycast plan: host tables");
  }
  group_views(views, nviews, ctx->quality, buf, &p->g);  // the quality is fixed at create
  // partial slots for the largest launch any tuning can make
  uint32_t slots = 0;
  for (int s = 0; s < NSHAPES; s++) {
    const r3d_raycast_tuning t{SHAPE_W[s], SHAPE_H[s], 0u, 0.0f};
    const uint32_t n = partial_slots(ctx, p->g, t);
    slots = n > slots ? n : slots;
  }
  const uint32_t pers = (uint32_t)(ctx->num_cus > 0 ? ctx->num_cus : 256) * MAX_BLOCKS_PER_CU * NKINDS;
  slots = pers > slots ? pers : slots;
  const size_t off_tab = (p->g.view_bytes() + 255u) & ~(size_t)255u;
  const size_t off_q = (off_tab + sizeof(uint32_t) * p->g.table_words() * NSHAPES + 255u) & ~(size_t)255u;
  const size_t off_part = off_q + 256u;
  const size_t off_order = (off_part + (size_t)slots * 32u + 255u) & ~(size_t)255u;
  const size_t total = off_order + order_words(p->g) * 4u;
  if (hipMalloc(&p->d_mem, total) != hipSuccess) {
    free(buf);
    delete p;
    return r3d_fail(ctx, R3D_ERR_NOMEM, "raycast plan: %zu device bytes", total);
  }
  p->d_views = (r3d_frame_params *)p->d_mem;
  p->d_tables = (uint32_t *)((uint8_t *)p->d_mem + off_tab);
  p->d_queue = (uint32_t *)((uint8_t *)p->d_mem + off_q);
  p->d_partials = (uint32_t *)((uint8_t *)p->d_mem + off_part);
  p->d_order = (uint32_t *)((uint8_t *)p->d_mem + off_order);
  hipError_t e = hipMemcpy(p->d_views, buf, p->g.view_bytes(), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(p->d_tables, buf + p->g.view_bytes(), sizeof(uint32_t) * p->g.table_words() * NSHAPES,
                  hipMemcpyHostToDevice);
  free(buf);
  if (e != hipSuccess) {
    hipFree(p->d_mem);
    delete p;
    return r3d_fail(ctx, R3D_ERR_DEVICE, "raycast plan: upload: %s", hipGetErrorString(e));
  }
  *plan = p;
  return R3D_OK;
}

extern "C" int r3d_hip_raycast_plan_launch(r3d_hip_raycast_plan *p, void *stream) {
  if (!p) return R3D_ERR_INVALID;
  r3d_hip_ctx *ctx = p->ctx;
  R3D_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  r3d_panes_reset(ctx);
  return launch_groups(ctx, base_args(ctx, p->vol, &p->out, &p->opts), p->d_views, p->d_tables, p->g, p->tuning,
                       p->d_partials, p->d_queue, p->d_order, (unsigned long long *)p->opts.counters,
                       4u, (hipStream_t)stream);
}

extern "C" int r3d_hip_raycast_plan_set_tuning(r3d_hip_raycast_plan *p, const r3d_raycast_tuning *t) {
  if (!p) return R3D_ERR_INVALID;
  const int rc = check_tuning(p->ctx, t);
  if (rc != R3D_OK) return rc;
  p->tuning = *t;
  return R3D_OK;
}

extern "C" int r3d_hip_raycast_plan_get_tuning(const r3d_hip_raycast_plan *p, r3d_raycast_tuning *t) {
  if (!p || !t) return R3D_ERR_INVALID;
  *t = p->tuning;
  return R3D_OK;
}

// Times every tile shape at one workgroup per tile and at 2, 4, 8 and 16
// persistent workgroups per CU (one warm-up launch, then the mean of three),
// keeps the fastest in the plan and returns the measurements.
extern "C" int r3d_hip_raycast_plan_tune(r3d_hip_raycast_plan *p, void *stream, r3d_raycast_tuning *results,
                                         uint32_t max_results, uint32_t *nresults) {
  if (!p) return R3D_ERR_INVALID;
  r3d_hip_ctx *ctx = p->ctx;
  R3D_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  hipStream_t s = (hipStream_t)stream;
  static const uint32_t bpc[5] = {0u, 2u, 4u, 8u, 16u};
  hipEvent_t e0, e1;
  R3D_HIP_CHECK(ctx, hipEventCreate(&e0));
  if (hipEventCreate(&e1) != hipSuccess) {
    hipEventDestroy(e0);
    return r3d_fail(ctx, R3D_ERR_DEVICE, "raycast tune: event");
  }
  const RayArgs a = base_args(ctx, p->vol, &p->out, &p->opts);
  r3d_raycast_tuning best{8u, 8u, 0u, 3.0e38f};
  uint32_t n = 0;
  int rc = R3D_OK;
  for (int sh = 0; sh < NSHAPES && rc == R3D_OK; sh++)
    for (int b = 0; b < 5 && rc == R3D_OK; b++) {
      r3d_raycast_tuning t{SHAPE_W[sh], SHAPE_H[sh], bpc[b], 0.0f};
      rc = launch_groups(ctx, a, p->d_views, p->d_tables, p->g, t, p->d_partials, p->d_queue, p->d_order, nullptr,
                         4u, s);
      if (rc != R3D_OK) break;
      if (hipEventRecord(e0, s) != hipSuccess) rc = R3D_ERR_DEVICE;
      for (int r = 0; r < 3 && rc == R3D_OK; r++)
        rc = launch_groups(ctx, a, p->d_views, p->d_tables, p->g, t, p->d_partials, p->d_queue, p->d_order, nullptr,
                         4u, s);
      float ms = 0.0f;
      if (rc == R3D_OK && (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                           hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
        rc = r3d_fail(ctx, R3D_ERR_DEVICE, "raycast tune: timing");
      if (rc != R3D_OK) break;
      t.ms = ms / 3.0f;
      if (results && n < max_results) results[n] = t;
      n++;
      if (t.ms < best.ms) best = t;
    }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (rc != R3D_OK) return rc;
  p->tuning = best;
  if (nresults) *nresults = n;
  return R3D_OK;
}

extern "C" void r3d_hip_raycast_plan_destroy(r3d_hip_raycast_plan *p) {
  if (!p) return;
  hipSetDevice(p->ctx->device);
  hipDeviceSynchronize();
  hipFree(p->d_mem);
  delete p;
}
