// r3d_raycast.hip -- cube-mode raycaster for gfx950.
//
// One lane per ray, 8x8-pixel tiles as 64-lane workgroups (one wave), every
// view of a call in one launch per mode class (tile -> view through a prefix
// table). Per lane the march is src/shaders/raycast.slang main() with
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
#include <string.h>

#include "r3d_ray.h"

namespace r3d {

struct RayArgs {
  VolDev vol;
  const r3d_frame_params *views;  // this launch's views
  const uint32_t *tile_start;     // nviews + 1 prefix sums of tiles
  uint32_t nviews;
  const float *tf;
  float *img;
  uint8_t *img8;
  uint32_t img_w, img_h;
  float cutoff;
  uint32_t *partials;  // [tile][4] counter sums of this launch's tiles, or NULL
  uint32_t *px_steps;
};

struct RayOut {
  float c[4];
  uint32_t steps, skips, shaded;
};

// The per-ray march. KIND fixes the compositing / shading branches at compile
// time, as the reference builds one pipeline per mode (src/shaders/shaders.cmake:17-30).
template <int KIND>
__device__ __forceinline__ RayOut march(const Sampler &S, const VolDev &V, const float4 *tf,
                                        const r3d_frame_params &pc, float cutoff, uint32_t px,
                                        uint32_t py) {
  RayOut R;
  R.steps = R.skips = R.shaded = 0;
  const uint32_t mode = KIND == KIND_FULL ? MODE_FULL : KIND == KIND_MIP ? MODE_MIP : KIND == KIND_FLAT ? MODE_FLAT : pc.mode;
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
  const float voxel = 1.0f / (float)V.maxdim;
  const float max_lod = fmin_sel(pc.max_mip, (float)(V.nlev - 1u));
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
  const float span_step = fmax_sel((tfar - tn) / ((float)MAX_STEPS * dt), 1.0f);
  const float bs = 8.0f * voxel;
  const float occ_gate = fmax_sel(pc.skip_gate, 0.5f / 255.0f);
  const float opac_k = pc.density * pc.step_voxels;
  const V3 lvec = vscale(dir, -1.0f);
  const V3 hvec = vnorm(vsub(lvec, dir));
  const V3 key = mk(0x1.a20bd8p-2f, 0x1.a20bd8p-1f, -0x1.a20bd8p-2f);  // normalize(0.4,0.8,-0.4)
  const float sgx = dir.x > 0.0f ? 1.0f : 0.0f, sgy = dir.y > 0.0f ? 1.0f : 0.0f,
              sgz = dir.z > 0.0f ? 1.0f : 0.0f;
  while (t < tfar && steps < MAX_STEPS) {
    steps++;
    const float foot = ortho ? cone : t * cone;
    const float lod = fclamp_sel(log2_pinned(fmax_sel(foot / voxel, 1.0f)) + pc.lod_bias, 0.0f, max_lod);
    const float lod_step = fmax_sel(exp2_pinned(lod), span_step);
    const V3 p = vadd(ro, vscale(dir, t));
    float v = S.sample(p, lod);
    const bool cut = v < pc.threshold;
    if (cut) v = 0.0f;
    if (v < occ_gate && occ_sample(V, p) < occ_gate) {  // raycast.slang:866-869
      const float nx = (__builtin_floorf(p.x / bs) + sgx) * bs;
      const float ny = (__builtin_floorf(p.y / bs) + sgy) * bs;
      const float nz = (__builtin_floorf(p.z / bs) + sgz) * bs;
      const float ex = __builtin_fabsf(dir.x) > 1e-8f ? (nx - ro.x) / dir.x : 1e30f;
      const float ey = __builtin_fabsf(dir.y) > 1e-8f ? (ny - ro.y) / dir.y : 1e30f;
      const float ez = __builtin_fabsf(dir.z) > 1e-8f ? (nz - ro.z) / dir.z : 1e30f;
      t = fmax_sel(fmin_sel(fmin_sel(ex, ey), ez), t) + bs * 0.01f;
      skips++;
      continue;
    }
    if (v > mip_max) mip_max = v;
    if (mode == MODE_MIP && mip_max >= 0.999f) break;
    if (KIND == KIND_OTHER && first_hit < 0.0f && v > 0.05f) first_hit = t;
    if (composite && !cut) {
      const float4 s = tf_sample(tf, v);
      const float a = 1.0f - pow_unit(1.0f - s.w, opac_k * lod_step);
      if (a > 0.003f) {
        float cr = s.x, cg = s.y, cb = s.z;
        if (shaded) {
          shaded_n++;
          const float e = voxel * lod_step;
          const float gx = S.sample(mk(p.x + e, p.y, p.z), lod) - S.sample(mk(p.x - e, p.y, p.z), lod);
          const float gy = S.sample(mk(p.x, p.y + e, p.z), lod) - S.sample(mk(p.x, p.y - e, p.z), lod);
          const float gz = S.sample(mk(p.x, p.y, p.z + e), lod) - S.sample(mk(p.x, p.y, p.z - e), lod);
          const V3 g = mk(gx, gy, gz);
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
        const float wa = (1.0f - aa) * a;
        ar = ar + wa * cr;
        ag = ag + wa * cg;
        ab = ab + wa * cb;
        aa = aa + wa;
        if (aa >= cutoff) break;
      }
    }
    t += dt * lod_step;
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

template <int KIND>
__global__ __launch_bounds__(64) void k_raycast(RayArgs a) {
  __shared__ float4 s_tf[256];
  __shared__ LevelDev s_lev[R3D_MAX_LEVELS];
  const uint32_t lane = threadIdx.x;
  for (uint32_t i = lane; i < 256u; i += 64u)
    s_tf[i] = make_float4(a.tf[i * 4 + 0], a.tf[i * 4 + 1], a.tf[i * 4 + 2], a.tf[i * 4 + 3]);
  if (lane < a.vol.nlev) s_lev[lane] = a.vol.levels[lane];
  __syncthreads();
  // tile -> view
  const uint32_t tile = blockIdx.x;
  uint32_t v = 0;
  while (v + 1u < a.nviews && a.tile_start[v + 1] <= tile) v++;
  const r3d_frame_params &pc = a.views[v];
  const bool half = (pc.view_flags & 32u) != 0u;
  const uint32_t W = pc.viewport[0], H = pc.viewport[1];
  const uint32_t TW = half ? (W + 1u) / 2u : W, TH = half ? (H + 1u) / 2u : H;
  const uint32_t tiles_x = (TW + TILE - 1u) / TILE;
  const uint32_t lt = tile - a.tile_start[v];
  const uint32_t tx = (lt % tiles_x) * TILE + (lane % TILE);
  const uint32_t ty = (lt / tiles_x) * TILE + (lane / TILE);
  uint32_t st = 0, sk = 0, sh = 0, nr = 0;
  if (tx < TW && ty < TH) {
    const uint32_t lx = half ? tx * 2u : tx, ly = half ? ty * 2u : ty;
    Sampler S{a.vol.base, s_lev, a.vol.nlev};
    const RayOut R = march<KIND>(S, a.vol, s_tf, pc, a.cutoff, lx, ly);
    st = R.steps;
    sk = R.skips;
    sh = R.shaded;
    nr = 1;
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
  if (a.partials) {  // one plain store per tile; k_reduce_counters sums them
    st = wave_sum(st);
    sk = wave_sum(sk);
    sh = wave_sum(sh);
    nr = wave_sum(nr);
    if (lane == 0) ((uint4 *)a.partials)[tile] = make_uint4(st, sk, sh, nr);
  }
}

__global__ __launch_bounds__(256) void k_reduce_counters(const uint32_t *partials, uint32_t ntiles,
                                                         unsigned long long *counters) {
  __shared__ unsigned long long red[4][256];
  unsigned long long acc[4] = {0, 0, 0, 0};
  for (uint32_t i = threadIdx.x; i < ntiles; i += 256u) {
    const uint4 p = ((const uint4 *)partials)[i];
    acc[0] += p.x;
    acc[1] += p.y;
    acc[2] += p.z;
    acc[3] += p.w;
  }
  for (int c = 0; c < 4; c++) red[c][threadIdx.x] = acc[c];
  __syncthreads();
  for (uint32_t s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s)
      for (int c = 0; c < 4; c++) red[c][threadIdx.x] += red[c][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x < 4) counters[threadIdx.x] += red[threadIdx.x][0];
}

__global__ void k_zero_u64(unsigned long long *p, uint32_t n) {
  if (threadIdx.x < n) p[threadIdx.x] = 0ull;
}

static int kind_of(uint32_t mode) {
  switch (mode) {
  case MODE_FULL: return KIND_FULL;
  case MODE_MIP: return KIND_MIP;
  case MODE_FLAT: return KIND_FLAT;
  default: return KIND_OTHER;
  }
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

extern "C" int r3d_hip_raycast(r3d_hip_ctx *ctx, const r3d_hip_volume *vol, const r3d_frame_params *views,
                               uint32_t nviews, const r3d_hip_image *out, const r3d_raycast_opts *opts,
                               r3d_hip_stats *stats, void *stream) {
  if (!ctx) return R3D_ERR_INVALID;
  if (!vol || vol->ctx != ctx) return r3d_fail(ctx, R3D_ERR_INVALID, "raycast: bad volume");
  if (!views || nviews == 0 || nviews > R3D_MAX_VIEWS_PER_CALL)
    return r3d_fail(ctx, R3D_ERR_INVALID, "raycast: nviews %u outside 1..%u", nviews, R3D_MAX_VIEWS_PER_CALL);
  if (!out || (!out->rgba && !out->rgba8) || out->width == 0 || out->height == 0)
    return r3d_fail(ctx, R3D_ERR_INVALID, "raycast: no output image");
  if (out->width > 65535u || out->height > 65535u)
    return r3d_fail(ctx, R3D_ERR_RANGE, "raycast: image above 65535 pixels per side");
  const float cutoff = opts && opts->alpha_cutoff > 0.0f ? opts->alpha_cutoff : 0.98f;
  for (uint32_t i = 0; i < nviews; i++) {
    const r3d_frame_params &p = views[i];
    if (p.slab_grid || p.clip_valid || p.brick_mode || (p.view_flags & 2u))
      return r3d_fail(ctx, R3D_ERR_UNSUPPORTED, "raycast: view %u is not a cube-mode view", i);
    if (p.mode > MODE_FLAT) return r3d_fail(ctx, R3D_ERR_INVALID, "raycast: view %u mode %u", i, p.mode);
    if (p.viewport[0] == 0 || p.viewport[1] == 0 || p.viewport[0] > 65535u || p.viewport[1] > 65535u)
      return r3d_fail(ctx, R3D_ERR_INVALID, "raycast: view %u viewport %ux%u", i, p.viewport[0], p.viewport[1]);
    if (!(p.step_voxels > 0.0f)) return r3d_fail(ctx, R3D_ERR_INVALID, "raycast: view %u step_voxels", i);
    if (!(p.max_mip >= 0.0f)) return r3d_fail(ctx, R3D_ERR_INVALID, "raycast: view %u max_mip", i);
  }
  R3D_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  hipStream_t s = (hipStream_t)stream;

  // group views by kernel kind; staging layout: [views by group][tile starts]
  const size_t pbytes = sizeof(r3d_frame_params) * nviews;
  const size_t tbytes = sizeof(uint32_t) * (nviews + NKINDS);
  uint8_t buf[sizeof(r3d_frame_params) * R3D_MAX_VIEWS_PER_CALL + sizeof(uint32_t) * (R3D_MAX_VIEWS_PER_CALL + NKINDS)];
  r3d_frame_params *pv = (r3d_frame_params *)buf;
  uint32_t *ts = (uint32_t *)(buf + pbytes);
  uint32_t gstart[NKINDS + 1], gtiles[NKINDS], tpos[NKINDS];
  uint32_t n = 0, tcur = 0;
  for (int k = 0; k < NKINDS; k++) {
    gstart[k] = n;
    tpos[k] = tcur;
    uint32_t tiles = 0;
    for (uint32_t i = 0; i < nviews; i++) {
      if (kind_of(views[i].mode) != k) continue;
      pv[n] = views[i];
      ts[tcur++] = tiles;
      const bool half = (views[i].view_flags & 32u) != 0u;
      const uint32_t TW = half ? (views[i].viewport[0] + 1u) / 2u : views[i].viewport[0];
      const uint32_t TH = half ? (views[i].viewport[1] + 1u) / 2u : views[i].viewport[1];
      tiles += ((TW + TILE - 1u) / TILE) * ((TH + TILE - 1u) / TILE);
      n++;
    }
    ts[tcur++] = tiles;
    gtiles[k] = tiles;
  }
  gstart[NKINDS] = n;
  ViewSlot *slot = nullptr;
  int rc = stage(ctx, buf, pbytes + tbytes, s, &slot);
  if (rc != R3D_OK) return rc;
  unsigned long long *counters = opts ? (unsigned long long *)opts->counters : nullptr;
  if (stats) {
    counters = ctx->d_counters;
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, s, counters, 4u);
  }
  const uint32_t total_tiles = gtiles[0] + gtiles[1] + gtiles[2] + gtiles[3];
  if (counters && ctx->partials_cap < total_tiles) {
    if (ctx->d_partials) {
      R3D_HIP_CHECK(ctx, hipStreamSynchronize(s));
      R3D_HIP_CHECK(ctx, hipFree(ctx->d_partials));
      ctx->d_partials = nullptr;
      ctx->partials_cap = 0;
    }
    const size_t cap = total_tiles < 65536u ? 65536u : total_tiles;
    if (hipMalloc(&ctx->d_partials, cap * 16u) != hipSuccess)
      return r3d_fail(ctx, R3D_ERR_NOMEM, "raycast: counter partials for %zu tiles", cap);
    ctx->partials_cap = cap;
  }
  RayArgs a;
  a.vol = vol_dev(vol);
  a.tf = opts && opts->tf ? opts->tf : ctx->d_tf;
  a.img = out->rgba;
  a.img8 = out->rgba8;
  a.img_w = out->width;
  a.img_h = out->height;
  a.cutoff = cutoff;
  a.px_steps = opts ? opts->px_steps : nullptr;
  const r3d_frame_params *dviews = (const r3d_frame_params *)slot->dev;
  const uint32_t *dts = (const uint32_t *)((const uint8_t *)slot->dev + pbytes);
  uint32_t tile_base = 0;
  for (int k = 0; k < NKINDS; k++) {
    const uint32_t nv = gstart[k + 1] - gstart[k];
    if (nv == 0 || gtiles[k] == 0) continue;
    a.partials = counters ? ctx->d_partials + (size_t)tile_base * 4u : nullptr;
    tile_base += gtiles[k];
    a.views = dviews + gstart[k];
    a.tile_start = dts + tpos[k];
    a.nviews = nv;
    const dim3 grid(gtiles[k]), block(64);
    switch (k) {
    case KIND_FULL: hipLaunchKernelGGL(k_raycast<KIND_FULL>, grid, block, 0, s, a); break;
    case KIND_MIP: hipLaunchKernelGGL(k_raycast<KIND_MIP>, grid, block, 0, s, a); break;
    case KIND_FLAT: hipLaunchKernelGGL(k_raycast<KIND_FLAT>, grid, block, 0, s, a); break;
    default: hipLaunchKernelGGL(k_raycast<KIND_OTHER>, grid, block, 0, s, a); break;
    }
    R3D_HIP_CHECK(ctx, hipGetLastError());
  }
  if (counters && total_tiles) {
    hipLaunchKernelGGL(k_reduce_counters, dim3(1), dim3(256), 0, s, (const uint32_t *)ctx->d_partials,
                       total_tiles, counters);
    R3D_HIP_CHECK(ctx, hipGetLastError());
  }
  R3D_HIP_CHECK(ctx, hipEventRecord(slot->ev, s));
  slot->recorded = true;
  if (stats) {
    unsigned long long h[4];
    R3D_HIP_CHECK(ctx, hipMemcpyAsync(h, counters, sizeof h, hipMemcpyDeviceToHost, s));
    R3D_HIP_CHECK(ctx, hipStreamSynchronize(s));
    stats->samples = h[0];
    stats->skips = h[1];
    stats->shaded = h[2];
    stats->rays = h[3];
  }
  return R3D_OK;
}
