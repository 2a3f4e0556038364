/* oracle_raycast.c -- TEST INFRASTRUCTURE ONLY (the oracle; see r3d_oracle.h).
 *
 * Host transcription of the cube-mode raycaster, src/shaders/raycast.slang
 * main() (lines 505-1039) specialised to R3D_MODE 0, over a linear mip
 * pyramid. The texture unit's filtering is replaced by the pinned FP32
 * sequence below (the hardware's sub-texel rounding is not reproducible):
 *
 *   tri(level, p):  t = clamp(p * dim - 0.5, 0, dim - 1) per axis,
 *                   i0 = (int)t, i1 = min(i0 + 1, dim - 1), f = t - i0;
 *                   c00 = lerp(V[z0][y0][x0], V[z0][y0][x1], fx) ... (values
 *                   are the raw u8 as floats), then lerp over y, then z
 *                   (the order of tests/test_gpu.c:33-49).
 *   sample(p, lod): d = (int)lod, w = lod - d; s = tri(d); if (w > 0 and
 *                   d + 1 < nlev) s = lerp(s, tri(d + 1), w); return s * K255
 *                   (LINEAR mip filter, src/vk/vkbackend.c:1058-1069).
 *   tf(v):          t = clamp(v * 256 - 0.5, 0, 255), i0 = (int)t,
 *                   i1 = min(i0 + 1, 255), lerp per channel
 *                   (tests/test_gpu.c:51-56).
 *   occ(p):         nearest texel floor(p * odim) clamped, value * K255.
 *   lerp(a,b,t) = fmaf(t, b - a, a); dot = (ax*bx + ay*by) + az*bz unfused;
 *   length = sqrtf(dot); normalize(v) = v / length(v); min/max are selects.
 *   K255 = 1.0f / 255.0f rounded once.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle_math.h"
#include "r3d_oracle.h"

#define K255 (1.0f / 255.0f)
#define MAX_STEPS 4096u      /* raycast.slang:120 */
#define OCC_BLOCK_F 8.0f     /* raycast.slang:125 */

enum { MODE_FULL = 0, MODE_MIP = 1, MODE_DEPTH = 2, MODE_HEATMAP = 3, MODE_RAYDIR = 4, MODE_FLAT = 5 };

typedef struct { float x, y, z; } v3f;
static inline v3f mk(float x, float y, float z) { v3f r = {x, y, z}; return r; }
static inline v3f vadd(v3f a, v3f b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3f vsub(v3f a, v3f b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3f vscale(v3f a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
static inline float vdot(v3f a, v3f b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline float vlen(v3f a) { return sqrtf(vdot(a, a)); }
static inline v3f vnorm(v3f a) { float l = vlen(a); return mk(a.x / l, a.y / l, a.z / l); }

/* ---- planted mistakes ----------------------------------------------------
 * The tests' sensitivity checks switch one plausible mistake on (r3do_plant)
 * and require the closed-form check it targets to fail; 0 (the default) is
 * the oracle as transcribed. */
static int g_plant = R3DO_PLANT_NONE;
void r3do_plant(int k) { g_plant = k; }

/* ---- sampling ---------------------------------------------------------- */

static float tri(const r3do_volume *V, uint32_t l, v3f p) {
  const uint32_t nx = V->dims[l][0], ny = V->dims[l][1], nz = V->dims[l][2];
  float tx = om_clamp(p.x * (float)nx - 0.5f, 0.0f, (float)(nx - 1));
  float ty = om_clamp(p.y * (float)ny - 0.5f, 0.0f, (float)(ny - 1));
  float tz = om_clamp(p.z * (float)nz - 0.5f, 0.0f, (float)(nz - 1));
  uint32_t x0 = (uint32_t)tx, y0 = (uint32_t)ty, z0 = (uint32_t)tz;
  uint32_t x1 = x0 + 1 < nx ? x0 + 1 : nx - 1;
  uint32_t y1 = y0 + 1 < ny ? y0 + 1 : ny - 1;
  uint32_t z1 = z0 + 1 < nz ? z0 + 1 : nz - 1;
  float fx = tx - (float)x0, fy = ty - (float)y0, fz = tz - (float)z0;
  if (V->touched) { /* the footprint's brick: level l's bricks follow those of levels 0..l-1 */
    uint64_t first = 0;
    for (uint32_t k = 0; k < l; k++)
      first += (uint64_t)((V->dims[k][0] + 15) / 16) * ((V->dims[k][1] + 15) / 16) * ((V->dims[k][2] + 15) / 16);
    const uint64_t gb = first + ((uint64_t)(z0 >> 4) * ((ny + 15) / 16) + (y0 >> 4)) * ((nx + 15) / 16) + (x0 >> 4);
    __atomic_fetch_or(&V->touched[gb >> 5], 1u << (gb & 31u), __ATOMIC_RELAXED);
  }
  const uint8_t *L = V->lev[l];
#define VV(Z, Y, X) ((float)L[((size_t)(Z) * ny + (Y)) * nx + (X)])
  float c00 = om_lerp(VV(z0, y0, x0), VV(z0, y0, x1), fx);
  float c10 = om_lerp(VV(z0, y1, x0), VV(z0, y1, x1), fx);
  float c01 = om_lerp(VV(z1, y0, x0), VV(z1, y0, x1), fx);
  float c11 = om_lerp(VV(z1, y1, x0), VV(z1, y1, x1), fx);
#undef VV
  return om_lerp(om_lerp(c00, c10, fy), om_lerp(c01, c11, fy), fz);
}

/* vol_sample for cube mode: volumes[0].SampleLevel(p, lod) (raycast.slang:448). */
static float vol_sample(const r3do_volume *V, v3f p, float lod) {
  uint32_t d = (uint32_t)lod;
  float w = lod - (float)d;
  float s = tri(V, d, p);
  if (w > 0.0f && d + 1 < V->nlev) s = om_lerp(s, tri(V, d + 1, p), g_plant == R3DO_PLANT_MIP_WEIGHT ? 1.0f - w : w);
  return s * K255;
}

static float occ_sample(const r3do_volume *V, v3f p) {
  float q[3] = {p.x, p.y, p.z};
  uint32_t idx[3];
  for (int a = 0; a < 3; a++) {
    float f = floorf(q[a] * (float)V->odim[a]);
    float hi = (float)(V->odim[a] - 1);
    f = om_clamp(f, 0.0f, hi);
    idx[a] = (uint32_t)f;
  }
  uint8_t m = V->occ[((size_t)idx[2] * V->odim[1] + idx[1]) * V->odim[0] + idx[0]];
  return (float)m * K255;
}

static void tf_sample(const float *tf, float v, float out[4]) {
  float t = om_clamp(v * 256.0f - 0.5f, 0.0f, 255.0f);
  uint32_t i0 = (uint32_t)t;
  uint32_t i1 = i0 + 1 < 256 ? i0 + 1 : 255;
  float f = t - (float)i0;
  for (int c = 0; c < 4; c++) out[c] = om_lerp(tf[i0 * 4 + c], tf[i1 * 4 + c], f);
}

/* ---- gradient shading (raycast.slang:915-999) -------------------------- */

#define IR3 0.57735026919f    /* 1/sqrt(3): tetrahedron tap radius (raycast.slang:926) */
#define RENORM 0.86602540378f /* sqrt(3)/2 (raycast.slang:927) */

/* Four tetrahedron-corner taps q0..q3 at p + ee*(+,-,-), (-,-,+), (-,+,-),
 * (+,+,+) to a gradient (raycast.slang:952-953), left to right in FP32. */
static v3f tet_grad(float q0, float q1, float q2, float q3) {
  const float r = g_plant == R3DO_PLANT_FAST_RENORM ? 1.0f : RENORM;
  return mk((q0 - q1 - q2 + q3) * r, (-q0 - q1 + q2 + q3) * r, (-q0 + q1 - q2 + q3) * r);
}

/* Cube mode: six central-difference taps at radius e, or (fast) the four
 * tetrahedral taps at radius e/sqrt(3) (raycast.slang:946-951, 991-994). */
static v3f grad_cube(const r3do_volume *V, v3f p, float lod, float e, int fast) {
  if (fast) {
    const float ee = e * IR3;
    const float q0 = vol_sample(V, mk(p.x + ee, p.y + -ee, p.z + -ee), lod);
    const float q1 = vol_sample(V, mk(p.x + -ee, p.y + -ee, p.z + ee), lod);
    const float q2 = vol_sample(V, mk(p.x + -ee, p.y + ee, p.z + -ee), lod);
    const float q3 = vol_sample(V, mk(p.x + ee, p.y + ee, p.z + ee), lod);
    return tet_grad(q0, q1, q2, q3);
  }
  return mk(vol_sample(V, mk(p.x + e, p.y, p.z), lod) - vol_sample(V, mk(p.x - e, p.y, p.z), lod),
            vol_sample(V, mk(p.x, p.y + e, p.z), lod) - vol_sample(V, mk(p.x, p.y - e, p.z), lod),
            vol_sample(V, mk(p.x, p.y, p.z + e), lod) - vol_sample(V, mk(p.x, p.y, p.z - e), lod));
}

/* Gradient-modulated Blinn-Phong (raycast.slang:985-997): headlight l = -dir,
 * key light normalize(0.4, 0.8, -0.4) rounded once, half vector
 * h = normalize(l - dir); pow(x, 24) by squaring. */
static void shade_col(v3f g, v3f lvec, v3f hvec, float col[3]) {
  const v3f key = mk(0x1.a20bd8p-2f, 0x1.a20bd8p-1f, -0x1.a20bd8p-2f);
  const float gm = vlen(g);
  if (!(gm > 1e-4f)) return;
  const v3f n = mk(g.x / gm, g.y / gm, g.z / gm);
  const float kd = g_plant == R3DO_PLANT_KEY_LIGHT ? 0.0f : 0.3f * fabsf(vdot(n, key));
  const float diff = 0.7f * fabsf(vdot(n, lvec)) + kd;
  const float nh = fabsf(vdot(n, hvec));
  const float nh2 = nh * nh, nh4 = nh2 * nh2, nh8 = nh4 * nh4;
  const float spec = (g_plant == R3DO_PLANT_SPEC_EXP ? nh8 * nh8 : om_pow24(nh)) * 0.3f;
  const float lit = 0.25f + 0.75f * diff;
  const float w = om_sat(gm * 6.0f);
  const float m = g_plant == R3DO_PLANT_LIT_LERP ? om_lerp(lit, 1.0f, w) : om_lerp(1.0f, lit, w);
  const float sw = spec * w;
  col[0] = col[0] * m + sw;
  col[1] = col[1] * m + sw;
  col[2] = col[2] * m + sw;
}

/* ---- ray setup (raycast.slang:469-495, common.slang:45-52) ------------- */

static v3f to_volume_dir(const r3do_frame_params *pc, v3f d) {
  return mk(vdot(mk(pc->vol_r0[0], pc->vol_r1[0], pc->vol_r2[0]), d),
            vdot(mk(pc->vol_r0[1], pc->vol_r1[1], pc->vol_r2[1]), d),
            vdot(mk(pc->vol_r0[2], pc->vol_r1[2], pc->vol_r2[2]), d));
}
static v3f to_volume_point(const r3do_frame_params *pc, v3f p, v3f c) {
  v3f q = vsub(vsub(p, mk(pc->vol_tx, pc->vol_ty, pc->vol_tz)), c);
  return vadd(to_volume_dir(pc, q), c);
}

static void pix_ndc(const r3do_frame_params *pc, uint32_t px, uint32_t py, float *nx, float *ny) {
  float u = ((float)px + 0.5f) / (float)pc->viewport[0];
  float v = ((float)py + 0.5f) / (float)pc->viewport[1];
  *nx = u * 2.0f - 1.0f;
  *ny = 1.0f - v * 2.0f;
}

/* ---- one pixel ---------------------------------------------------------- */

static void march(const r3do_volume *V, const float *tf, const r3do_frame_params *pc,
                  float cutoff, int fast, uint32_t px, uint32_t py, float out[4], uint32_t *nsteps,
                  uint64_t *nskip, uint64_t *nshade, uint64_t *nblend, uint64_t *nblend_sh) {
  const v3f fwd = mk(pc->cam_forward[0], pc->cam_forward[1], pc->cam_forward[2]);
  const v3f right = mk(pc->cam_right[0], pc->cam_right[1], pc->cam_right[2]);
  const v3f up = mk(pc->cam_up[0], pc->cam_up[1], pc->cam_up[2]);
  const int ortho = (pc->view_flags & 1u) != 0u;
  float ndx, ndy;
  pix_ndc(pc, px, py, &ndx, &ndy);
  v3f dir = ortho ? vnorm(fwd) : vnorm(vadd(vadd(fwd, vscale(right, ndx)), vscale(up, ndy)));
  *nsteps = 0;
  if (pc->mode == MODE_RAYDIR) {
    out[0] = dir.x * 0.5f + 0.5f; out[1] = dir.y * 0.5f + 0.5f; out[2] = dir.z * 0.5f + 0.5f;
    out[3] = 1.0f;
    return;
  }
  uint32_t md = V->dims[0][0];
  if (V->dims[0][1] > md) md = V->dims[0][1];
  if (V->dims[0][2] > md) md = V->dims[0][2];
  const float voxel = 1.0f / (float)md;
  const float max_lod = om_min(pc->max_mip, (float)(V->nlev - 1));
  const v3f ext = mk(1.0f, 1.0f, 1.0f);
  v3f org = mk(pc->cam_origin[0], pc->cam_origin[1], pc->cam_origin[2]);
  if (ortho) org = vadd(vadd(org, vscale(right, ndx)), vscale(up, ndy));
  const v3f ro = to_volume_point(pc, org, vscale(ext, 0.5f));
  dir = to_volume_dir(pc, dir);

  /* box_hit against [0, ext] (raycast.slang:489-495) */
  v3f inv = mk(1.0f / dir.x, 1.0f / dir.y, 1.0f / dir.z);
  v3f t0 = mk((0.0f - ro.x) * inv.x, (0.0f - ro.y) * inv.y, (0.0f - ro.z) * inv.z);
  v3f t1 = mk((ext.x - ro.x) * inv.x, (ext.y - ro.y) * inv.y, (ext.z - ro.z) * inv.z);
  v3f tmin = mk(om_min(t0.x, t1.x), om_min(t0.y, t1.y), om_min(t0.z, t1.z));
  v3f tmax = mk(om_max(t0.x, t1.x), om_max(t0.y, t1.y), om_max(t0.z, t1.z));
  float hx = om_max(om_max(tmin.x, tmin.y), tmin.z);
  float hy = om_min(om_min(tmax.x, tmax.y), tmax.z);
  float tn = om_max(hx, 0.0f), tfar = om_min(hy, 1e30f);
  if (tn >= tfar) {
    out[0] = 0.02f; out[1] = 0.02f; out[2] = 0.03f; out[3] = 1.0f;
    return;
  }
  /* raycast.slang:733-765 */
  const float cone = 2.0f * vlen(up) / (float)pc->viewport[1];
  const float dt = pc->step_voxels * voxel;
  const float jitter = om_hash12(px, py, pc->frame_index);
  float t = tn + jitter * dt;
  float mip_max = 0.0f, first_hit = -1.0f;
  float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
  uint32_t steps = 0;
  const int composite = pc->mode == MODE_FULL || pc->mode == MODE_FLAT;
  const int shaded = pc->mode == MODE_FULL;
  const float span_step = om_max((tfar - tn) / ((float)MAX_STEPS * dt), 1.0f);
  const float bs = OCC_BLOCK_F * voxel;
  const float occ_gate = om_max(pc->skip_gate, 0.5f / 255.0f);
  const float opac_k = pc->density * pc->step_voxels;
  const v3f lvec = vscale(dir, -1.0f);
  const v3f hvec = vnorm(vsub(lvec, dir));
  const v3f sgn = mk(dir.x > 0.0f ? 1.0f : 0.0f, dir.y > 0.0f ? 1.0f : 0.0f, dir.z > 0.0f ? 1.0f : 0.0f);

  while (t < tfar && steps < MAX_STEPS) {
    steps++;
    float foot = ortho ? cone : t * cone;
    float lod = om_clamp(om_log2f(om_max(foot / voxel, 1.0f)) + pc->lod_bias, 0.0f, max_lod);
    float lod_step = om_max(om_exp2f(lod), span_step);
    v3f p = vadd(ro, vscale(dir, t));
    float v = vol_sample(V, p, lod);
    const int blend = lod - (float)(uint32_t)lod > 0.0f && (uint32_t)lod + 1 < V->nlev;
    *nblend += (uint64_t)blend;
    int cut = v < pc->threshold;
    if (cut) v = 0.0f;
    /* empty-space skip (raycast.slang:866-869) */
    if (v < occ_gate && occ_sample(V, p) < occ_gate) {
      const v3f sg = g_plant == R3DO_PLANT_SKIP_FACE ? mk(0.0f, 0.0f, 0.0f) : sgn;
      v3f nextp = mk((floorf(p.x / bs) + sg.x) * bs, (floorf(p.y / bs) + sg.y) * bs,
                     (floorf(p.z / bs) + sg.z) * bs);
      float tx = fabsf(dir.x) > 1e-8f ? (nextp.x - ro.x) / dir.x : 1e30f;
      float ty = fabsf(dir.y) > 1e-8f ? (nextp.y - ro.y) / dir.y : 1e30f;
      float tz = fabsf(dir.z) > 1e-8f ? (nextp.z - ro.z) / dir.z : 1e30f;
      t = om_max(om_min(om_min(tx, ty), tz), t) + (g_plant == R3DO_PLANT_SKIP_NUDGE ? 0.0f : bs * 0.01f);
      (*nskip)++;
      continue;
    }
    if (v > mip_max) mip_max = v;
    if (pc->mode == MODE_MIP && mip_max >= 0.999f) break;
    if (first_hit < 0.0f && v > 0.05f) first_hit = t;
    if (composite && !cut) {
      float s[4];
      tf_sample(tf, v, s);
      float a = 1.0f - om_pow_unit(1.0f - s[3], opac_k * lod_step);
      if (a > 0.003f) {
        float col[3] = {s[0], s[1], s[2]};
        if (shaded) {
          (*nshade)++;
          *nblend_sh += (uint64_t)blend;
          shade_col(grad_cube(V, p, lod, g_plant == R3DO_PLANT_TAP_RADIUS ? voxel : voxel * lod_step, fast), lvec,
                    hvec, col);
        }
        float wa = (1.0f - acc[3]) * a;
        acc[0] = acc[0] + wa * col[0];
        acc[1] = acc[1] + wa * col[1];
        acc[2] = acc[2] + wa * col[2];
        acc[3] = acc[3] + wa;
        if (acc[3] >= cutoff) break;
      }
    }
    t += dt * lod_step;
  }
  *nsteps = steps;
  if (pc->mode == MODE_HEATMAP) {
    float x = (float)steps / 1024.0f;
    out[0] = x; out[1] = 1.0f - fabsf(x - 0.5f) * 2.0f; out[2] = 1.0f - x;
  } else if (pc->mode == MODE_DEPTH) {
    float d = first_hit < 0.0f ? 0.0f : 1.0f - om_sat(first_hit / 2.0f);
    out[0] = d; out[1] = d; out[2] = d;
  } else if (composite) {
    float ta = 1.0f - acc[3];
    out[0] = acc[0] + ta * 0.02f; out[1] = acc[1] + ta * 0.02f; out[2] = acc[2] + ta * 0.03f;
  } else {
    float m = mip_max * pc->density;
    out[0] = m; out[1] = m; out[2] = m;
  }
  out[3] = 1.0f;
}

static int view_supported(const r3do_frame_params *pc) {
  /* cube mode only: slab / clip / bricks / virtual-slab / surf views are the
   * other compile-time specialisations (vkbackend.c:7125-7130) */
  if (pc->slab_grid || pc->clip_valid || pc->brick_mode) return 0;
  if (pc->view_flags & 2u) return 0;
  if (pc->mode > MODE_FLAT) return 0;
  if (pc->viewport[0] == 0 || pc->viewport[1] == 0) return 0;
  if (!(pc->step_voxels > 0.0f) || !(pc->max_mip >= 0.0f)) return 0;
  return 1;
}

int r3do_raycast(const r3do_volume *vol, const float *tf, const r3do_frame_params *views,
                 uint32_t nviews, float alpha_cutoff, float *img, uint32_t img_w, uint32_t img_h,
                 uint32_t *px_steps, uint32_t row0, uint32_t row_step, int nthreads,
                 r3do_counters *cnt, int quality) {
  for (uint32_t v = 0; v < nviews; v++)
    if (!view_supported(&views[v])) return -1;
  if (row_step == 0) row_step = 1;
  if (nthreads < 1) nthreads = 1;
  uint64_t tot_s = 0, tot_k = 0, tot_h = 0, tot_r = 0, tot_b = 0, tot_bh = 0;
  for (uint32_t v = 0; v < nviews; v++) {
    const r3do_frame_params *pc = &views[v];
    const int half = (pc->view_flags & 32u) != 0u;
    const uint32_t ox = pc->view_org & 0xffffu, oy = pc->view_org >> 16;
    const uint32_t W = pc->viewport[0], H = pc->viewport[1];
    const uint32_t TW = half ? (W + 1) / 2 : W, TH = half ? (H + 1) / 2 : H;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads) \
    reduction(+ : tot_s, tot_k, tot_h, tot_r, tot_b, tot_bh)
    for (int64_t ty = 0; ty < (int64_t)TH; ty++) {
      if ((uint32_t)ty % row_step != row0 % row_step) continue;
      for (uint32_t tx = 0; tx < TW; tx++) {
        uint32_t lx = half ? tx * 2 : tx, ly = half ? (uint32_t)ty * 2 : (uint32_t)ty;
        if (lx >= W || ly >= H) continue;
        float c[4];
        uint32_t ns = 0;
        uint64_t nk = 0, nh = 0, nb = 0, nbh = 0;
        march(vol, tf, pc, alpha_cutoff, quality == 1, lx, ly, c, &ns, &nk, &nh, &nb, &nbh);
        tot_s += ns; tot_k += nk; tot_h += nh; tot_r += 1; tot_b += nb; tot_bh += nbh;
        for (uint32_t dy = 0; dy < (half ? 2u : 1u); dy++)
          for (uint32_t dx = 0; dx < (half ? 2u : 1u); dx++) {
            uint32_t qx = lx + dx, qy = ly + dy;
            if (qx >= W || qy >= H) continue;
            uint32_t gx = ox + qx, gy = oy + qy;
            if (gx >= img_w || gy >= img_h) continue;
            memcpy(img + ((size_t)gy * img_w + gx) * 4, c, sizeof c);
            if (px_steps) px_steps[(size_t)gy * img_w + gx] = ns;
          }
      }
    }
  }
  if (cnt) {
    cnt->samples = tot_s; cnt->skips = tot_k; cnt->shaded = tot_h; cnt->rays = tot_r;
    cnt->blended = tot_b; cnt->blended_shaded = tot_bh;
  }
  return 0;
}

/* ---- bricks-LOD mode ----------------------------------------------------
 * raycast.slang main() with R3D_MODE 3 and BR_LOD_FLAG (lines 142-279 and the
 * bricks branches of 640-1039), overlays off. The page table is the
 * reference's (vkbackend.c:4185-4205): page[0] = nlev, page[1..3] = base shape
 * (x, y, z), per level at 4 + 4l {offset of its entries, packed brick dims
 * bx | by << 10 | bz << 20, scale}; an entry is slot | max << 24, ~0 when the
 * brick is absent. Slots are 128^3 u8, slot s at atlas + s * 2 MiB. The
 * texture fetch is the pinned trilinear of the cube path inside the slot:
 *   t = clamp(bl, 0.5, 127.5) - 0.5 per axis, i0 = (int)t, i1 = min(i0+1, 127),
 *   lerp over x, then y, then z, times K255 (bricks_fetch, atlas lod 0). */
#define BR_SLOT 128.0f
#define BR_SLOT_VOX (128u * 128u * 128u)

typedef struct {
  int ok;
  float maxv;
  v3f bl;
  uint32_t slot;
  float grad_vox, world_brick;
} bctx;

typedef struct {
  int valid, desired;
  uint32_t scale, bc[3];
  bctx c;
} lodcache;

static float brick_fetch(const r3do_bricks *B, uint32_t slot, v3f bl) {
  const uint8_t *S = B->atlas + (size_t)slot * BR_SLOT_VOX;
  const float tx = om_clamp(bl.x, 0.5f, 127.5f) - 0.5f;
  const float ty = om_clamp(bl.y, 0.5f, 127.5f) - 0.5f;
  const float tz = om_clamp(bl.z, 0.5f, 127.5f) - 0.5f;
  const uint32_t x0 = (uint32_t)tx, y0 = (uint32_t)ty, z0 = (uint32_t)tz;
  const uint32_t x1 = x0 + 1 < 128 ? x0 + 1 : 127, y1 = y0 + 1 < 128 ? y0 + 1 : 127, z1 = z0 + 1 < 128 ? z0 + 1 : 127;
  const float fx = tx - (float)x0, fy = ty - (float)y0, fz = tz - (float)z0;
#define BV(Z, Y, X) ((float)S[((size_t)(Z) * 128u + (Y)) * 128u + (X)])
  const float c00 = om_lerp(BV(z0, y0, x0), BV(z0, y0, x1), fx);
  const float c10 = om_lerp(BV(z0, y1, x0), BV(z0, y1, x1), fx);
  const float c01 = om_lerp(BV(z1, y0, x0), BV(z1, y0, x1), fx);
  const float c11 = om_lerp(BV(z1, y1, x0), BV(z1, y1, x1), fx);
#undef BV
  return om_lerp(om_lerp(c00, c10, fy), om_lerp(c01, c11, fy), fz) * K255;
}

/* shared by the locate and its cache: the sample position in base voxels */
static v3f base_voxel(const uint32_t *page, v3f p, float *maxdim) {
  const float sx = (float)page[1], sy = (float)page[2], sz = (float)page[3];
  const float md = om_max(sx, om_max(sy, sz));
  *maxdim = md;
  return mk(om_clamp(p.x * md, 0.0f, om_max(sx - 0.001f, 0.0f)), om_clamp(p.y * md, 0.0f, om_max(sy - 0.001f, 0.0f)),
            om_clamp(p.z * md, 0.0f, om_max(sz - 0.001f, 0.0f)));
}

/* bricks_lod_locate_ex (raycast.slang:198-247): desired level, +1, +2, then the coarsest */
static bctx locate_ex(const r3do_bricks *B, v3f p, float lod, uint32_t *scale_out, uint32_t bc_out[3]) {
  const uint32_t *page = B->page;
  bctx c;
  memset(&c, 0, sizeof c);
  *scale_out = 1;
  bc_out[0] = bc_out[1] = bc_out[2] = 0;
  c.grad_vox = 1.0f;
  const uint32_t nlev = page[0];
  float maxdim;
  const v3f bv = base_voxel(page, p, &maxdim);
  c.world_brick = BR_SLOT * (float)(1u << (nlev - 1u)) / maxdim;
  const int desired = (int)om_clamp(floorf(lod), 0.0f, (float)(nlev - 1u));
  const int last = (int)nlev - 1;
  for (uint32_t step = 0; step < 4u; step++) {
    const int li = step < 3u ? desired + (int)step : last;
    if (li < 0 || li > last || (step == 3u && li <= desired + 2)) continue;
    const uint32_t h = 4u + (uint32_t)li * 4u;
    const uint32_t off = page[h], packed = page[h + 1u], scale = page[h + 2u];
    const uint32_t bdx = packed & 1023u, bdy = (packed >> 10) & 1023u, bdz = packed >> 20;
    const v3f lv = mk(bv.x / (float)scale, bv.y / (float)scale, bv.z / (float)scale);
    uint32_t bcx = (uint32_t)lv.x >> 7, bcy = (uint32_t)lv.y >> 7, bcz = (uint32_t)lv.z >> 7;
    if (bcx > bdx - 1u) bcx = bdx - 1u;
    if (bcy > bdy - 1u) bcy = bdy - 1u;
    if (bcz > bdz - 1u) bcz = bdz - 1u;
    const uint32_t e = page[off + (bcz * bdy + bcy) * bdx + bcx];
    if (e == 0xFFFFFFFFu) continue;
    c.ok = 1;
    c.maxv = (float)(e >> 24) / 255.0f;
    c.slot = e & 0xFFFFFFu;
    c.bl = mk(lv.x - (float)(bcx << 7), lv.y - (float)(bcy << 7), lv.z - (float)(bcz << 7));
    c.grad_vox = om_exp2f(lod) / (float)scale;
    c.world_brick = BR_SLOT * (float)scale / maxdim;
    *scale_out = scale;
    bc_out[0] = bcx;
    bc_out[1] = bcy;
    bc_out[2] = bcz;
    return c;
  }
  return c;
}

/* bricks_lod_locate_cached (raycast.slang:255-279) */
static bctx locate_cached(const r3do_bricks *B, v3f p, float lod, lodcache *lc) {
  const uint32_t nlev = B->page[0];
  const int desired = (int)om_clamp(floorf(lod), 0.0f, (float)(nlev - 1u));
  if (lc->valid && lc->desired == desired) {
    float maxdim;
    const v3f bv = base_voxel(B->page, p, &maxdim);
    const float sc = (float)lc->scale;
    const v3f lv = mk(bv.x / sc, bv.y / sc, bv.z / sc);
    const uint32_t bcx = (uint32_t)lv.x >> 7, bcy = (uint32_t)lv.y >> 7, bcz = (uint32_t)lv.z >> 7;
    if (bcx == lc->bc[0] && bcy == lc->bc[1] && bcz == lc->bc[2]) {
      bctx c = lc->c;
      c.bl = mk(lv.x - (float)(bcx << 7), lv.y - (float)(bcy << 7), lv.z - (float)(bcz << 7));
      return c;
    }
  }
  uint32_t scale, bc[3];
  bctx c = locate_ex(B, p, lod, &scale, bc);
  lc->valid = c.ok;
  lc->desired = desired;
  lc->scale = scale;
  lc->bc[0] = bc[0];
  lc->bc[1] = bc[1];
  lc->bc[2] = bc[2];
  lc->c = c;
  return c;
}

/* Bricks-LOD taps inside the located brick, in brick-local voxels
 * (raycast.slang:933-937, 964-973). */
static v3f grad_brick(const r3do_bricks *B, uint32_t slot, v3f b, float ev, int fast) {
  if (fast) {
    const float e = ev * IR3;
    const float q0 = brick_fetch(B, slot, mk(b.x + e, b.y + -e, b.z + -e));
    const float q1 = brick_fetch(B, slot, mk(b.x + -e, b.y + -e, b.z + e));
    const float q2 = brick_fetch(B, slot, mk(b.x + -e, b.y + e, b.z + -e));
    const float q3 = brick_fetch(B, slot, mk(b.x + e, b.y + e, b.z + e));
    return tet_grad(q0, q1, q2, q3);
  }
  return mk(brick_fetch(B, slot, mk(b.x + ev, b.y, b.z)) - brick_fetch(B, slot, mk(b.x + -ev, b.y, b.z)),
            brick_fetch(B, slot, mk(b.x, b.y + ev, b.z)) - brick_fetch(B, slot, mk(b.x, b.y + -ev, b.z)),
            brick_fetch(B, slot, mk(b.x, b.y, b.z + ev)) - brick_fetch(B, slot, mk(b.x, b.y, b.z + -ev)));
}

static void march_bricks(const r3do_bricks *B, const float *tf, const r3do_frame_params *pc, float cutoff,
                         int fast, uint32_t px, uint32_t py, float out[4], uint32_t *nsteps, uint64_t *nskip,
                         uint64_t *nshade) {
  const uint32_t *page = B->page;
  const v3f fwd = mk(pc->cam_forward[0], pc->cam_forward[1], pc->cam_forward[2]);
  const v3f right = mk(pc->cam_right[0], pc->cam_right[1], pc->cam_right[2]);
  const v3f up = mk(pc->cam_up[0], pc->cam_up[1], pc->cam_up[2]);
  const int ortho = (pc->view_flags & 1u) != 0u;
  float ndx, ndy;
  pix_ndc(pc, px, py, &ndx, &ndy);
  v3f dir = ortho ? vnorm(fwd) : vnorm(vadd(vadd(fwd, vscale(right, ndx)), vscale(up, ndy)));
  *nsteps = 0;
  if (pc->mode == MODE_RAYDIR) {
    out[0] = dir.x * 0.5f + 0.5f; out[1] = dir.y * 0.5f + 0.5f; out[2] = dir.z * 0.5f + 0.5f;
    out[3] = 1.0f;
    return;
  }
  /* raycast.slang:655-660, 305-309 */
  const float sx = (float)page[1], sy = (float)page[2], sz = (float)page[3];
  const float md = (float)(page[1] > page[2] ? (page[1] > page[3] ? page[1] : page[3]) : (page[2] > page[3] ? page[2] : page[3]));
  const float voxel = 1.0f / md;
  const float max_lod = (float)(page[0] - 1u);
  const float emax = om_max(sx, om_max(sy, sz));
  const v3f ext = mk(sx / emax, sy / emax, sz / emax);
  v3f org = mk(pc->cam_origin[0], pc->cam_origin[1], pc->cam_origin[2]);
  if (ortho) org = vadd(vadd(org, vscale(right, ndx)), vscale(up, ndy));
  const v3f ro = to_volume_point(pc, org, vscale(ext, 0.5f));
  dir = to_volume_dir(pc, dir);
  /* raycast.slang:690-718: zoom-crop box or slab range */
  v3f lo = mk(0.0f, 0.0f, 0.0f), hi = ext;
  float t_lo = 0.0f, t_hi = 1e30f;
  if ((pc->view_flags & 16u) != 0u) {
    const v3f clo = mk(pc->slab_x0 / md, pc->slab_y0 / md, pc->slab_z0 / md);
    const v3f chi = vadd(clo, mk(pc->slab_px / md, pc->slab_py / md, pc->slab_nx / md));
    lo = mk(om_clamp(clo.x, 0.0f, ext.x), om_clamp(clo.y, 0.0f, ext.y), om_clamp(clo.z, 0.0f, ext.z));
    hi = mk(om_clamp(chi.x, lo.x, ext.x), om_clamp(chi.y, lo.y, ext.y), om_clamp(chi.z, lo.z, ext.z));
  } else if (pc->slab_depth > 0u) {
    if ((pc->view_flags & 4u) != 0u) {
      t_lo = pc->slab_z0 / md;
      t_hi = (pc->slab_z0 + (float)pc->slab_depth) / md;
    } else {
      const uint32_t axc = (pc->view_flags >> 8) & 3u;
      const int a = axc == 1u ? 0 : (axc == 2u ? 1 : 2);
      float *plo = a == 0 ? &lo.x : a == 1 ? &lo.y : &lo.z, *phi = a == 0 ? &hi.x : a == 1 ? &hi.y : &hi.z;
      const float ea = a == 0 ? ext.x : a == 1 ? ext.y : ext.z;
      *plo = om_clamp(pc->slab_z0 / md, 0.0f, ea);
      *phi = om_clamp((pc->slab_z0 + (float)pc->slab_depth) / md, *plo, ea);
    }
  }
  v3f inv = mk(1.0f / dir.x, 1.0f / dir.y, 1.0f / dir.z);
  v3f t0 = mk((lo.x - ro.x) * inv.x, (lo.y - ro.y) * inv.y, (lo.z - ro.z) * inv.z);
  v3f t1 = mk((hi.x - ro.x) * inv.x, (hi.y - ro.y) * inv.y, (hi.z - ro.z) * inv.z);
  float hx = om_max(om_max(om_min(t0.x, t1.x), om_min(t0.y, t1.y)), om_min(t0.z, t1.z));
  float hy = om_min(om_min(om_max(t0.x, t1.x), om_max(t0.y, t1.y)), om_max(t0.z, t1.z));
  float tn = om_max(hx, t_lo), tfar = om_min(hy, t_hi);
  if (tn >= tfar) {
    out[0] = 0.02f; out[1] = 0.02f; out[2] = 0.03f; out[3] = 1.0f;
    return;
  }
  const float cone = 2.0f * vlen(up) / (float)pc->viewport[1];
  const float dt = pc->step_voxels * voxel;
  float t = tn + om_hash12(px, py, pc->frame_index) * dt;
  float mip_max = 0.0f, first_hit = -1.0f;
  float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
  uint32_t steps = 0;
  const int composite = pc->mode == MODE_FULL || pc->mode == MODE_FLAT;
  const int shaded = pc->mode == MODE_FULL;
  const uint32_t budget = shaded ? 192u : 256u; /* raycast.slang:743 */
  const float span_step = om_max((tfar - tn) / ((float)budget * dt), 1.0f);
  const float occ_gate = om_max(pc->skip_gate, 0.5f / 255.0f);
  const float bgate = om_max(occ_gate, 4.5f / 255.0f);
  const float opac_k = pc->density * pc->step_voxels;
  const v3f lvec = vscale(dir, -1.0f);
  const v3f hvec = vnorm(vsub(lvec, dir));
  const v3f sgn = mk(dir.x > 0.0f ? 1.0f : 0.0f, dir.y > 0.0f ? 1.0f : 0.0f, dir.z > 0.0f ? 1.0f : 0.0f);
  lodcache lc;
  memset(&lc, 0, sizeof lc);
  lc.desired = -1;
  lc.scale = 1u;
  while (t < tfar && steps < budget) {
    steps++;
    const float foot = ortho ? cone : t * cone;
    const float lod = om_clamp(om_log2f(om_max(foot / voxel, 1.0f)) + pc->lod_bias, 0.0f, max_lod);
    const float lod_step = om_max(om_exp2f(lod), span_step);
    const v3f p = vadd(ro, vscale(dir, t));
    const bctx bc = locate_cached(B, p, lod, &lc);
    float v = bc.ok ? brick_fetch(B, bc.slot, bc.bl) : 0.0f;
    const int cut = v < pc->threshold;
    if (cut) v = 0.0f;
    /* raycast.slang:834-868: whole-brick skips */
    float bsz = 0.0f;
    if (pc->mode == MODE_MIP && bc.ok && bc.maxv <= mip_max)
      bsz = bc.world_brick;
    else if (v < bgate && bc.maxv < bgate)
      bsz = bc.world_brick;
    if (bsz > 0.0f) {
      v3f nextp = mk((floorf(p.x / bsz) + sgn.x) * bsz, (floorf(p.y / bsz) + sgn.y) * bsz,
                     (floorf(p.z / bsz) + sgn.z) * bsz);
      float tx = fabsf(dir.x) > 1e-8f ? (nextp.x - ro.x) / dir.x : 1e30f;
      float ty = fabsf(dir.y) > 1e-8f ? (nextp.y - ro.y) / dir.y : 1e30f;
      float tz = fabsf(dir.z) > 1e-8f ? (nextp.z - ro.z) / dir.z : 1e30f;
      t = om_max(om_min(om_min(tx, ty), tz), t) + bsz * 0.01f;
      (*nskip)++;
      continue;
    }
    if (v > mip_max) mip_max = v;
    if (pc->mode == MODE_MIP && mip_max >= 0.999f) break;
    if (first_hit < 0.0f && v > 0.05f) first_hit = t;
    if (composite && !cut) {
      float s[4];
      tf_sample(tf, v, s);
      float a = 1.0f - om_pow_unit(1.0f - s[3], opac_k * lod_step);
      if (a > 0.003f) {
        float col[3] = {s[0], s[1], s[2]};
        if (shaded) {
          (*nshade)++;
          /* taps reuse the located brick (raycast.slang:964-973) */
          shade_col(grad_brick(B, bc.slot, bc.bl, bc.grad_vox, fast), lvec, hvec, col);
        }
        float wa = (1.0f - acc[3]) * a;
        acc[0] = acc[0] + wa * col[0];
        acc[1] = acc[1] + wa * col[1];
        acc[2] = acc[2] + wa * col[2];
        acc[3] = acc[3] + wa;
        if (acc[3] >= cutoff) break;
      }
    }
    t += dt * lod_step;
  }
  *nsteps = steps;
  if (pc->mode == MODE_HEATMAP) {
    float x = (float)steps / 1024.0f;
    out[0] = x; out[1] = 1.0f - fabsf(x - 0.5f) * 2.0f; out[2] = 1.0f - x;
  } else if (pc->mode == MODE_DEPTH) {
    float d = first_hit < 0.0f ? 0.0f : 1.0f - om_sat(first_hit / 2.0f);
    out[0] = d; out[1] = d; out[2] = d;
  } else if (composite) {
    float ta = 1.0f - acc[3];
    out[0] = acc[0] + ta * 0.02f; out[1] = acc[1] + ta * 0.02f; out[2] = acc[2] + ta * 0.03f;
  } else {
    float m = mip_max * pc->density;
    out[0] = m; out[1] = m; out[2] = m;
  }
  out[3] = 1.0f;
}

int r3do_raycast_bricks(const r3do_bricks *B, const float *tf, const r3do_frame_params *views, uint32_t nviews,
                        float alpha_cutoff, float *img, uint32_t img_w, uint32_t img_h, uint32_t *px_steps,
                        uint32_t row0, uint32_t row_step, int nthreads, r3do_counters *cnt, int quality) {
  for (uint32_t v = 0; v < nviews; v++) {
    const r3do_frame_params *pc = &views[v];
    if (!(pc->brick_mode & 0x20000u) || (pc->brick_mode & 0xffu) != B->abpa) return -1;
    if (pc->slab_grid || pc->clip_valid || (pc->view_flags & 2u) || pc->mode > MODE_FLAT) return -1;
    if (pc->viewport[0] == 0 || pc->viewport[1] == 0 || !(pc->step_voxels > 0.0f)) return -1;
  }
  if (row_step == 0) row_step = 1;
  if (nthreads < 1) nthreads = 1;
  uint64_t tot_s = 0, tot_k = 0, tot_h = 0, tot_r = 0;
  for (uint32_t v = 0; v < nviews; v++) {
    const r3do_frame_params *pc = &views[v];
    const int half = (pc->view_flags & 32u) != 0u;
    const uint32_t ox = pc->view_org & 0xffffu, oy = pc->view_org >> 16;
    const uint32_t W = pc->viewport[0], H = pc->viewport[1];
    const uint32_t TW = half ? (W + 1) / 2 : W, TH = half ? (H + 1) / 2 : H;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads) reduction(+ : tot_s, tot_k, tot_h, tot_r)
    for (int64_t ty = 0; ty < (int64_t)TH; ty++) {
      if ((uint32_t)ty % row_step != row0 % row_step) continue;
      for (uint32_t tx = 0; tx < TW; tx++) {
        uint32_t lx = half ? tx * 2 : tx, ly = half ? (uint32_t)ty * 2 : (uint32_t)ty;
        if (lx >= W || ly >= H) continue;
        float c[4];
        uint32_t ns = 0;
        uint64_t nk = 0, nh = 0;
        march_bricks(B, tf, pc, alpha_cutoff, quality == 1, lx, ly, c, &ns, &nk, &nh);
        tot_s += ns; tot_k += nk; tot_h += nh; tot_r += 1;
        for (uint32_t dy = 0; dy < (half ? 2u : 1u); dy++)
          for (uint32_t dx = 0; dx < (half ? 2u : 1u); dx++) {
            uint32_t qx = lx + dx, qy = ly + dy;
            if (qx >= W || qy >= H) continue;
            uint32_t gx = ox + qx, gy = oy + qy;
            if (gx >= img_w || gy >= img_h) continue;
            memcpy(img + ((size_t)gy * img_w + gx) * 4, c, sizeof c);
            if (px_steps) px_steps[(size_t)gy * img_w + gx] = ns;
          }
      }
    }
  }
  if (cnt) {
    cnt->samples = tot_s; cnt->skips = tot_k; cnt->shaded = tot_h; cnt->rays = tot_r;
    cnt->blended = 0; cnt->blended_shaded = 0; /* one level per sample */
  }
  return 0;
}

/* The page table and atlas of a pyramid (levels as linear [z][y][x] arrays,
 * level l of scale 2^l): per level ceil(dim / 128) bricks per axis, entries
 * level by level after the 64-word header; a brick is resident unless
 * resident[] (one flag per entry, in page order, or NULL) says 0. Resident
 * bricks take slots in page order; a slot holds the brick's voxels and zeros
 * past the level's edge; its entry is slot | max << 24. Returns the number of
 * page words, or 0 when page_cap / nslots are too small. */
uint32_t r3do_bricks_build(const r3do_volume *V, const uint8_t *resident, uint32_t *page, uint32_t page_cap,
                           uint8_t *atlas, uint32_t nslots) {
  uint32_t words = 64u, offs[16];
  for (uint32_t l = 0; l < V->nlev; l++) {
    offs[l] = words;
    const uint32_t bx = (V->dims[l][0] + 127u) / 128u, by = (V->dims[l][1] + 127u) / 128u,
                   bz = (V->dims[l][2] + 127u) / 128u;
    words += bx * by * bz;
  }
  if (words > page_cap) return 0;
  memset(page, 0, 64u * sizeof *page);
  page[0] = V->nlev;
  page[1] = V->dims[0][0];
  page[2] = V->dims[0][1];
  page[3] = V->dims[0][2];
  uint32_t slot = 0;
  for (uint32_t l = 0; l < V->nlev; l++) {
    const uint32_t nx = V->dims[l][0], ny = V->dims[l][1], nz = V->dims[l][2];
    const uint32_t bx = (nx + 127u) / 128u, by = (ny + 127u) / 128u, bz = (nz + 127u) / 128u;
    page[4u + l * 4u] = offs[l];
    page[5u + l * 4u] = bx | (by << 10) | (bz << 20);
    page[6u + l * 4u] = 1u << l;
    for (uint32_t k = 0; k < bx * by * bz; k++) {
      const uint32_t w = offs[l] + k;
      if (resident && !resident[w - 64u]) {
        page[w] = 0xFFFFFFFFu;
        continue;
      }
      if (slot >= nslots) return 0;
      const uint32_t cx = k % bx, cy = (k / bx) % by, cz = k / (bx * by);
      uint8_t *S = atlas + (size_t)slot * BR_SLOT_VOX;
      uint32_t mx = 0;
      for (uint32_t z = 0; z < 128u; z++)
        for (uint32_t y = 0; y < 128u; y++)
          for (uint32_t x = 0; x < 128u; x++) {
            const uint32_t gx = cx * 128u + x, gy = cy * 128u + y, gz = cz * 128u + z;
            const uint8_t v = gx < nx && gy < ny && gz < nz ? V->lev[l][((size_t)gz * ny + gy) * nx + gx] : 0u;
            S[((size_t)z * 128u + y) * 128u + x] = v;
            if (v > mx) mx = v;
          }
      page[w] = slot | (mx << 24);
      slot++;
    }
  }
  return words;
}

/* ---- pyramid and occupancy --------------------------------------------- */

void r3do_downsample(const uint8_t *src, uint32_t nx, uint32_t ny, uint32_t nz, uint8_t *dst) {
  uint32_t mx = nx > 1 ? nx / 2 : 1, my = ny > 1 ? ny / 2 : 1, mz = nz > 1 ? nz / 2 : 1;
#pragma omp parallel for schedule(static)
  for (int64_t z = 0; z < (int64_t)mz; z++)
    for (uint32_t y = 0; y < my; y++)
      for (uint32_t x = 0; x < mx; x++) {
        uint32_t sum = 4u;
        for (uint32_t dz = 0; dz < 2; dz++)
          for (uint32_t dy = 0; dy < 2; dy++)
            for (uint32_t dx = 0; dx < 2; dx++) {
              uint32_t sx = 2 * x + dx, sy = 2 * y + dy, sz = 2 * (uint32_t)z + dz;
              if (sx >= nx) sx = nx - 1;
              if (sy >= ny) sy = ny - 1;
              if (sz >= nz) sz = nz - 1;
              sum += src[((size_t)sz * ny + sy) * nx + sx];
            }
        dst[((size_t)z * my + y) * mx + x] = (uint8_t)(sum / 8u);
      }
}

void r3do_occupancy(const uint8_t *vol, uint32_t nx, uint32_t ny, uint32_t nz, uint8_t *occ) {
  uint32_t ox = (nx + 7) / 8, oy = (ny + 7) / 8, oz = (nz + 7) / 8;
  size_t n = (size_t)ox * oy * oz;
  uint8_t *raw = malloc(n);
  if (!raw) return;
#pragma omp parallel for schedule(static)
  for (int64_t bz = 0; bz < (int64_t)oz; bz++)
    for (uint32_t by = 0; by < oy; by++)
      for (uint32_t bx = 0; bx < ox; bx++) {
        uint8_t m = 0;
        for (uint32_t z = (uint32_t)bz * 8; z < (uint32_t)bz * 8 + 8 && z < nz; z++)
          for (uint32_t y = by * 8; y < by * 8 + 8 && y < ny; y++)
            for (uint32_t x = bx * 8; x < bx * 8 + 8 && x < nx; x++) {
              uint8_t v = vol[((size_t)z * ny + y) * nx + x];
              if (v > m) m = v;
            }
        raw[((size_t)bz * oy + by) * ox + bx] = m;
      }
#pragma omp parallel for schedule(static)
  for (int64_t z = 0; z < (int64_t)oz; z++)
    for (uint32_t y = 0; y < oy; y++)
      for (uint32_t x = 0; x < ox; x++) {
        uint8_t m = 0;
        for (int dz = -1; dz <= 1; dz++)
          for (int dy = -1; dy <= 1; dy++)
            for (int dx = -1; dx <= 1; dx++) {
              int64_t zz = z + dz, yy = (int64_t)y + dy, xx = (int64_t)x + dx;
              if (zz < 0 || yy < 0 || xx < 0 || zz >= oz || yy >= oy || xx >= ox) continue;
              uint8_t v = raw[((size_t)zz * oy + (size_t)yy) * ox + (size_t)xx];
              if (v > m) m = v;
            }
        occ[((size_t)z * oy + y) * ox + x] = m;
      }
  free(raw);
}

/* ---- exported math ------------------------------------------------------ */

float r3do_log2f(float x) { return om_log2f(x); }
float r3do_exp2f(float x) { return om_exp2f(x); }
float r3do_pow_unit(float b, float y) { return om_pow_unit(b, y); }
float r3do_hash12(uint32_t px, uint32_t py, uint32_t frame) { return om_hash12(px, py, frame); }

void r3do_math_batch(const float *x, uint32_t n, uint32_t fn, float *out) {
  for (uint32_t i = 0; i < n; i++) {
    switch (fn) {
    case 0: out[i] = om_log2f(x[i]); break;
    case 1: out[i] = om_exp2f(x[i]); break;
    case 2: out[i] = om_pow_unit(x[i], 1.7f); break;
    case 3: out[i] = sqrtf(x[i]); break;
    case 4: out[i] = 1.0f / x[i]; break;
    case 5: out[i] = om_pow24(x[i]); break;
    default: { float s, c; om_sincos_turns(x[i], &s, &c); out[i] = fn == 6 ? s : c; } break;
    }
  }
}

/* ---- differentiable raycast (FP64 reference) ------------------------------------
 * The compositing integral with an FP32 transfer table tf and a density grid g
 * (see include/r3d_hip.h r3d_hip_raycast_diff): positions, lod, volume samples
 * and the shading factors follow the pinned FP32 sequence (they do not depend
 * on the parameters); the TF lookup, the grid sample, the opacity correction,
 * the 0.003 gate, compositing and the cutoff are evaluated in double. With
 * dimg != NULL the reverse pass accumulates dL/dtf and dL/dgrid (+=). */

typedef struct {
  float t, lod, lod_step, v;
  v3f p;
  float m, sw;   /* shading: col = s.rgb * m + sw */
  int cut;
} diff_sample;

static void shade_f(const r3do_volume *V, v3f p, float lod, float e, v3f lvec, v3f hvec, float *m, float *sw) {
  const v3f key = mk(0x1.a20bd8p-2f, 0x1.a20bd8p-1f, -0x1.a20bd8p-2f);
  float gx = vol_sample(V, mk(p.x + e, p.y, p.z), lod) - vol_sample(V, mk(p.x - e, p.y, p.z), lod);
  float gy = vol_sample(V, mk(p.x, p.y + e, p.z), lod) - vol_sample(V, mk(p.x, p.y - e, p.z), lod);
  float gz = vol_sample(V, mk(p.x, p.y, p.z + e), lod) - vol_sample(V, mk(p.x, p.y, p.z - e), lod);
  v3f g = mk(gx, gy, gz);
  float gm = vlen(g);
  *m = 1.0f;
  *sw = 0.0f;
  if (gm > 1e-4f) {
    v3f n = mk(g.x / gm, g.y / gm, g.z / gm);
    float diff = 0.7f * fabsf(vdot(n, lvec)) + 0.3f * fabsf(vdot(n, key));
    float spec = om_pow24(fabsf(vdot(n, hvec))) * 0.3f;
    float lit = 0.25f + 0.75f * diff;
    float w = om_sat(gm * 6.0f);
    *m = om_lerp(1.0f, lit, w);
    *sw = spec * w;
  }
}

static double grid_d(const double *G, const uint32_t gd[3], v3f p, uint32_t idx[8], double w[8]) {
  float q[3] = {p.x, p.y, p.z};
  uint32_t i0[3], i1[3];
  double f[3];
  for (int a = 0; a < 3; a++) {
    float t = om_clamp(q[a] * (float)gd[a] - 0.5f, 0.0f, (float)(gd[a] - 1));
    i0[a] = (uint32_t)t;
    i1[a] = i0[a] + 1 < gd[a] ? i0[a] + 1 : gd[a] - 1;
    f[a] = (double)(t - (float)i0[a]);
  }
  double s = 0.0;
  for (int c = 0; c < 8; c++) {
    uint32_t x = (c & 1) ? i1[0] : i0[0], y = (c & 2) ? i1[1] : i0[1], z = (c & 4) ? i1[2] : i0[2];
    idx[c] = (z * gd[1] + y) * gd[0] + x;
    w[c] = ((c & 1) ? f[0] : 1 - f[0]) * ((c & 2) ? f[1] : 1 - f[1]) * ((c & 4) ? f[2] : 1 - f[2]);
    s += w[c] * G[idx[c]];
  }
  return s;
}

static double sat_d(double x) { return x < 0 ? 0 : (x > 1 ? 1 : x); }

static void diff_pixel(const r3do_volume *V, const double *tf, const double *G, const uint32_t gd[3],
                       const r3do_frame_params *pc, double cutoff, uint32_t px, uint32_t py, double out[4],
                       const double *dout, double *dtf, double *dG, diff_sample *buf) {
  const v3f fwd = mk(pc->cam_forward[0], pc->cam_forward[1], pc->cam_forward[2]);
  const v3f right = mk(pc->cam_right[0], pc->cam_right[1], pc->cam_right[2]);
  const v3f up = mk(pc->cam_up[0], pc->cam_up[1], pc->cam_up[2]);
  const int ortho = (pc->view_flags & 1u) != 0u;
  const int shaded = pc->mode == MODE_FULL;
  float ndx, ndy;
  pix_ndc(pc, px, py, &ndx, &ndy);
  v3f dir = ortho ? vnorm(fwd) : vnorm(vadd(vadd(fwd, vscale(right, ndx)), vscale(up, ndy)));
  uint32_t md = V->dims[0][0];
  if (V->dims[0][1] > md) md = V->dims[0][1];
  if (V->dims[0][2] > md) md = V->dims[0][2];
  const float voxel = 1.0f / (float)md;
  const float max_lod = om_min(pc->max_mip, (float)(V->nlev - 1));
  v3f org = mk(pc->cam_origin[0], pc->cam_origin[1], pc->cam_origin[2]);
  if (ortho) org = vadd(vadd(org, vscale(right, ndx)), vscale(up, ndy));
  const v3f ro = to_volume_point(pc, org, mk(0.5f, 0.5f, 0.5f));
  dir = to_volume_dir(pc, dir);
  v3f inv = mk(1.0f / dir.x, 1.0f / dir.y, 1.0f / dir.z);
  v3f t0 = mk((0.0f - ro.x) * inv.x, (0.0f - ro.y) * inv.y, (0.0f - ro.z) * inv.z);
  v3f t1 = mk((1.0f - ro.x) * inv.x, (1.0f - ro.y) * inv.y, (1.0f - ro.z) * inv.z);
  float hx = om_max(om_max(om_min(t0.x, t1.x), om_min(t0.y, t1.y)), om_min(t0.z, t1.z));
  float hy = om_min(om_min(om_max(t0.x, t1.x), om_max(t0.y, t1.y)), om_max(t0.z, t1.z));
  float tn = om_max(hx, 0.0f), tfar = om_min(hy, 1e30f);
  out[0] = 0.02; out[1] = 0.02; out[2] = 0.03; out[3] = 0.0;
  if (tn >= tfar) return;
  const float cone = 2.0f * vlen(up) / (float)pc->viewport[1];
  const float dt = pc->step_voxels * voxel;
  float t = tn + om_hash12(px, py, pc->frame_index) * dt;
  const float span_step = om_max((tfar - tn) / ((float)MAX_STEPS * dt), 1.0f);
  const float opac_k = pc->density * pc->step_voxels;
  const v3f lvec = vscale(dir, -1.0f);
  const v3f hvec = vnorm(vsub(lvec, dir));
  double C[3] = {0, 0, 0}, A = 0.0;
  uint32_t n = 0, steps = 0;
  while (t < tfar && steps < MAX_STEPS) {
    steps++;
    diff_sample S;
    S.t = t;
    float foot = ortho ? cone : t * cone;
    S.lod = om_clamp(om_log2f(om_max(foot / voxel, 1.0f)) + pc->lod_bias, 0.0f, max_lod);
    S.lod_step = om_max(om_exp2f(S.lod), span_step);
    S.p = vadd(ro, vscale(dir, t));
    S.v = vol_sample(V, S.p, S.lod);
    S.cut = S.v < pc->threshold;
    S.m = 1.0f;
    S.sw = 0.0f;
    int done = 0;
    if (!S.cut) {
      float tt = om_clamp(S.v * 256.0f - 0.5f, 0.0f, 255.0f);
      uint32_t i0 = (uint32_t)tt, i1 = i0 + 1 < 256 ? i0 + 1 : 255;
      double f = (double)(tt - (float)i0);
      double s[4];
      for (int c = 0; c < 4; c++) s[c] = tf[i0 * 4 + c] * (1 - f) + tf[i1 * 4 + c] * f;
      uint32_t gi[8];
      double gw[8];
      double g = grid_d(G, gd, S.p, gi, gw);
      double k = (double)opac_k * (double)S.lod_step * g;
      double b = 1.0 - sat_d(s[3]);
      double a = b <= 0 ? 1.0 : (b >= 1 ? 0.0 : 1.0 - pow(b, k));
      if (a > 0.003) {
        if (shaded) shade_f(V, S.p, S.lod, voxel * S.lod_step, lvec, hvec, &S.m, &S.sw);
        double wa = (1.0 - A) * a;
        for (int c = 0; c < 3; c++) C[c] += wa * (s[c] * S.m + S.sw);
        A += wa;
        buf[n++] = S;
        if (A >= cutoff) done = 1;
      }
    }
    if (done) break;
    t += dt * S.lod_step;
  }
  out[0] = C[0] + (1 - A) * 0.02;
  out[1] = C[1] + (1 - A) * 0.02;
  out[2] = C[2] + (1 - A) * 0.03;
  out[3] = A;
  if (!dout) return;
  /* reverse: recompute T before each contributing sample forward, then sweep back */
  double *Tb = (double *)malloc(sizeof(double) * (n ? n : 1));
  double T = 1.0;
  for (uint32_t i = 0; i < n; i++) {
    const diff_sample *S = &buf[i];
    float tt = om_clamp(S->v * 256.0f - 0.5f, 0.0f, 255.0f);
    uint32_t i0 = (uint32_t)tt, i1 = i0 + 1 < 256 ? i0 + 1 : 255;
    double f = (double)(tt - (float)i0);
    double sa = tf[i0 * 4 + 3] * (1 - f) + tf[i1 * 4 + 3] * f;
    uint32_t gi[8];
    double gw[8];
    double g = grid_d(G, gd, S->p, gi, gw);
    double k = (double)opac_k * (double)S->lod_step * g, b = 1.0 - sat_d(sa);
    double a = b <= 0 ? 1.0 : (b >= 1 ? 0.0 : 1.0 - pow(b, k));
    Tb[i] = T;
    T *= (1 - a);
  }
  const double gr = dout[0], gg = dout[1], gb = dout[2], ga = dout[3];
  double dT = gr * 0.02 + gg * 0.02 + gb * 0.03 - ga;
  for (int64_t i = (int64_t)n - 1; i >= 0; i--) {
    const diff_sample *S = &buf[i];
    float tt = om_clamp(S->v * 256.0f - 0.5f, 0.0f, 255.0f);
    uint32_t i0 = (uint32_t)tt, i1 = i0 + 1 < 256 ? i0 + 1 : 255;
    double f = (double)(tt - (float)i0);
    double s[4];
    for (int c = 0; c < 4; c++) s[c] = tf[i0 * 4 + c] * (1 - f) + tf[i1 * 4 + c] * f;
    uint32_t gi[8];
    double gw[8];
    double g = grid_d(G, gd, S->p, gi, gw);
    double kbase = (double)opac_k * (double)S->lod_step;
    double k = kbase * g, b = 1.0 - sat_d(s[3]);
    double a = b <= 0 ? 1.0 : (b >= 1 ? 0.0 : 1.0 - pow(b, k));
    double col[3];
    for (int c = 0; c < 3; c++) col[c] = s[c] * S->m + S->sw;
    double Tp = Tb[i];
    double gc = gr * col[0] + gg * col[1] + gb * col[2];
    double da = gc * Tp - dT * Tp;
    dT = dT * (1 - a) + a * gc;
    double ds[4];
    ds[0] = gr * Tp * a * S->m;
    ds[1] = gg * Tp * a * S->m;
    ds[2] = gb * Tp * a * S->m;
    ds[3] = (s[3] > 0 && s[3] < 1) ? da * k * (1 - a) / b : 0.0;
    for (int c = 0; c < 4; c++) {
      dtf[i0 * 4 + c] += (1 - f) * ds[c];
      dtf[i1 * 4 + c] += f * ds[c];
    }
    if (b > 0 && b < 1) {
      double dg = da * (-(1 - a) * log(b)) * kbase;
      for (int c = 0; c < 8; c++) dG[gi[c]] += gw[c] * dg;
    }
  }
  free(Tb);
}

int r3do_raycast_diff_window(const r3do_volume *vol, const double *tf, const double *grid, const uint32_t gd[3],
                             const r3do_frame_params *views, uint32_t nviews, double cutoff, double *img,
                             uint32_t img_w, uint32_t img_h, const double *dimg, double *dtf, double *dgrid,
                             uint32_t x0, uint32_t y0, uint32_t x1, uint32_t y1) {
  for (uint32_t v = 0; v < nviews; v++) {
    if (!view_supported(&views[v])) return -1;
    if (views[v].mode != MODE_FULL && views[v].mode != MODE_FLAT) return -1;
  }
  diff_sample *buf = (diff_sample *)malloc(sizeof(diff_sample) * MAX_STEPS);
  if (!buf) return -1;
  for (uint32_t v = 0; v < nviews; v++) {
    const r3do_frame_params *pc = &views[v];
    const int half = (pc->view_flags & 32u) != 0u;
    const uint32_t ox = pc->view_org & 0xffffu, oy = pc->view_org >> 16;
    const uint32_t W = pc->viewport[0], H = pc->viewport[1];
    const uint32_t TW = half ? (W + 1) / 2 : W, TH = half ? (H + 1) / 2 : H;
    for (uint32_t ty = y0; ty < TH && ty < y1; ty++)
      for (uint32_t tx = x0; tx < TW && tx < x1; tx++) {
        uint32_t lx = half ? tx * 2 : tx, ly = half ? ty * 2 : ty;
        double dout[4] = {0, 0, 0, 0}, c[4];
        if (dimg)
          for (uint32_t dy = 0; dy < (half ? 2u : 1u); dy++)
            for (uint32_t dx = 0; dx < (half ? 2u : 1u); dx++) {
              uint32_t qx = lx + dx, qy = ly + dy;
              if (qx >= W || qy >= H || ox + qx >= img_w || oy + qy >= img_h) continue;
              for (int k = 0; k < 4; k++) dout[k] += dimg[((size_t)(oy + qy) * img_w + ox + qx) * 4 + k];
            }
        diff_pixel(vol, tf, grid, gd, pc, cutoff, lx, ly, c, dimg ? dout : NULL, dtf, dgrid, buf);
        for (uint32_t dy = 0; dy < (half ? 2u : 1u); dy++)
          for (uint32_t dx = 0; dx < (half ? 2u : 1u); dx++) {
            uint32_t qx = lx + dx, qy = ly + dy;
            if (qx >= W || qy >= H || ox + qx >= img_w || oy + qy >= img_h) continue;
            for (int k = 0; k < 4; k++) img[((size_t)(oy + qy) * img_w + ox + qx) * 4 + k] = c[k];
          }
      }
  }
  free(buf);
  return 0;
}

int r3do_raycast_diff(const r3do_volume *vol, const double *tf, const double *grid, const uint32_t gd[3],
                      const r3do_frame_params *views, uint32_t nviews, double cutoff, double *img,
                      uint32_t img_w, uint32_t img_h, const double *dimg, double *dtf, double *dgrid) {
  return r3do_raycast_diff_window(vol, tf, grid, gd, views, nviews, cutoff, img, img_w, img_h, dimg, dtf, dgrid, 0,
                                  0, 0xffffffffu, 0xffffffffu);
}
