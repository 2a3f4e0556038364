// r3d_adjoint.hip -- differentiable cube-mode raycast and its hand-written
// adjoint with respect to the FP32 transfer table and a density-correction
// grid.
//
// The reference has no differentiable path; this follows its compositing
// integral (src/shaders/raycast.slang:893-1038) with two parameter sets:
//   s   = lerp(tf[i0], tf[i1], f)          (the TF lookup, 256 x RGBA FP32)
//   g   = trilinear(grid, p)               (density correction, init 1.0)
//   k   = density * step * lod_step * g
//   a   = 1 - (1 - sat(s.a))^k             (opacity correction, scaled by g)
//   if a > 0.003: col = s.rgb * m + sw     (m, sw: shading, volume-only)
//                 acc += (1 - acc_a) * a * col,  acc_a += (1 - acc_a) * a
//   out = (acc + (1 - acc_a) * bg, acc_a)
// Empty-space skipping is off in this mode (its gate depends on the TF). The
// forward stores, per ray, the loop state (t, acc_a) every 64 iterations; the
// adjoint walks each ray's segments last to first, recomputes the segment
// forward from its checkpoint (keeping t and the transmittance before every
// iteration), then runs the segment in reverse:
//   dC = dL/d rgb (constant along the ray), dT_N = dC . bg - dL/d alpha
//   per contributing sample (T = transmittance before it):
//     d col = dC * T * a;  d a = dC . (T col) - dT * T;  dT <- dT (1 - a) + a dC . col
//     d s.rgb = d col * m;  d s.a = d a * k (1 - a) / b  (0 < s.a < 1, b = 1 - s.a)
//     d g = d a * (-(1 - a) ln b) * density * step * lod_step
//   scattered to tf[i0], tf[i1] with weights (1 - f, f) and to the 8 grid
//   corners with the trilinear weights.
// TF gradients are summed per workgroup in LDS and written as per-block
// partials, then reduced in a fixed order; grid gradients are summed in
// registers while a ray stays in one grid cell, then added with float atomics.
#include <string.h>

#include "r3d_ray.h"

namespace r3d {

constexpr uint32_t CKPT = 64u;                   // iterations per checkpoint segment
constexpr uint32_t NCKPT = MAX_STEPS / CKPT;     // 64
struct RayRecord {
  uint32_t nsteps;
  uint32_t pad[3];
  float2 ck[NCKPT];  // (t, acc_a) before iteration j * CKPT
};
static_assert(sizeof(RayRecord) == 528, "ray record layout");

struct DiffArgs {
  VolDev vol;
  const r3d_frame_params *views;
  const uint32_t *tile_start;  // nviews + 1
  const uint32_t *ray_start;   // nviews + 1
  uint32_t nviews;
  uint32_t ntiles;
  const float *tf;             // 256 x 4
  const float *grid;           // gz x gy x gx
  uint32_t gx, gy, gz;
  float cutoff;
  float *img;                  // forward output
  const float *dimg;           // adjoint input dL/d rgba
  uint32_t img_w, img_h;
  RayRecord *rec;
  float *gtf_part;             // adjoint: [block][1024]
  float *dgrid;                // adjoint: accumulated in place
};

struct Ray {
  bool hit;
  uint32_t px, py;
  V3 ro, dir, lvec, hvec;
  float tn, tfar, voxel, max_lod, cone, dt, t0, span_step, opac_k;
  bool ortho;
};

// Ray setup exactly as the forward kernel (raycast.slang:640-765, cube mode).
__device__ __forceinline__ Ray setup_ray(const VolDev &V, const r3d_frame_params &pc, uint32_t px, uint32_t py) {
  Ray R;
  R.px = px;
  R.py = py;
  const V3 fwd = mk(pc.cam_forward[0], pc.cam_forward[1], pc.cam_forward[2]);
  const V3 right = mk(pc.cam_right[0], pc.cam_right[1], pc.cam_right[2]);
  const V3 up = mk(pc.cam_up[0], pc.cam_up[1], pc.cam_up[2]);
  R.ortho = (pc.view_flags & 1u) != 0u;
  const float u = ((float)px + 0.5f) / (float)pc.viewport[0];
  const float vv = ((float)py + 0.5f) / (float)pc.viewport[1];
  const float ndx = u * 2.0f - 1.0f, ndy = 1.0f - vv * 2.0f;
  V3 dir = R.ortho ? vnorm(fwd) : vnorm(vadd(vadd(fwd, vscale(right, ndx)), vscale(up, ndy)));
  R.voxel = 1.0f / (float)V.maxdim;
  R.max_lod = fmin_sel(pc.max_mip, (float)(V.nlev - 1u));
  V3 org = mk(pc.cam_origin[0], pc.cam_origin[1], pc.cam_origin[2]);
  if (R.ortho) org = vadd(vadd(org, vscale(right, ndx)), vscale(up, ndy));
  const V3 ctr = mk(0.5f, 0.5f, 0.5f);
  const V3 c0 = mk(pc.vol_r0[0], pc.vol_r1[0], pc.vol_r2[0]);
  const V3 c1 = mk(pc.vol_r0[1], pc.vol_r1[1], pc.vol_r2[1]);
  const V3 c2 = mk(pc.vol_r0[2], pc.vol_r1[2], pc.vol_r2[2]);
  const V3 q = vsub(vsub(org, mk(pc.vol_tx, pc.vol_ty, pc.vol_tz)), ctr);
  R.ro = vadd(mk(vdot(c0, q), vdot(c1, q), vdot(c2, q)), ctr);
  dir = mk(vdot(c0, dir), vdot(c1, dir), vdot(c2, dir));
  R.dir = dir;
  const V3 inv = mk(1.0f / dir.x, 1.0f / dir.y, 1.0f / dir.z);
  const V3 t0 = mk((0.0f - R.ro.x) * inv.x, (0.0f - R.ro.y) * inv.y, (0.0f - R.ro.z) * inv.z);
  const V3 t1 = mk((1.0f - R.ro.x) * inv.x, (1.0f - R.ro.y) * inv.y, (1.0f - R.ro.z) * inv.z);
  const float hx = fmax_sel(fmax_sel(fmin_sel(t0.x, t1.x), fmin_sel(t0.y, t1.y)), fmin_sel(t0.z, t1.z));
  const float hy = fmin_sel(fmin_sel(fmax_sel(t0.x, t1.x), fmax_sel(t0.y, t1.y)), fmax_sel(t0.z, t1.z));
  R.tn = fmax_sel(hx, 0.0f);
  R.tfar = fmin_sel(hy, 1e30f);
  R.hit = !(R.tn >= R.tfar);
  R.cone = 2.0f * vlen(up) / (float)pc.viewport[1];
  R.dt = pc.step_voxels * R.voxel;
  R.t0 = R.tn + hash12(px, py, pc.frame_index) * R.dt;
  R.span_step = fmax_sel((R.tfar - R.tn) / ((float)MAX_STEPS * R.dt), 1.0f);
  R.opac_k = pc.density * pc.step_voxels;
  R.lvec = vscale(dir, -1.0f);
  R.hvec = vnorm(vsub(R.lvec, dir));
  return R;
}

struct GridTap {
  float g;
  uint32_t i000;           // index of corner (x0,y0,z0)
  uint32_t dx, dy, dz;     // index steps to the +1 corners (0 at the edge)
  float fx, fy, fz;
};

// trilinear density correction, cell-centred, clamp-to-edge: the volume
// sampler's sequence on an FP32 grid
__device__ __forceinline__ GridTap grid_tap(const float *G, uint32_t gx, uint32_t gy, uint32_t gz, V3 p) {
  GridTap T;
  const float tx = fclamp_ord(p.x * (float)gx - 0.5f, 0.0f, (float)(gx - 1u));
  const float ty = fclamp_ord(p.y * (float)gy - 0.5f, 0.0f, (float)(gy - 1u));
  const float tz = fclamp_ord(p.z * (float)gz - 0.5f, 0.0f, (float)(gz - 1u));
  const uint32_t x0 = (uint32_t)tx, y0 = (uint32_t)ty, z0 = (uint32_t)tz;
  T.fx = tx - (float)x0;
  T.fy = ty - (float)y0;
  T.fz = tz - (float)z0;
  T.dx = x0 + 1u < gx ? 1u : 0u;
  T.dy = y0 + 1u < gy ? gx : 0u;
  T.dz = z0 + 1u < gz ? gx * gy : 0u;
  T.i000 = (z0 * gy + y0) * gx + x0;
  const float *q = G + T.i000;
  const float c00 = flerp(q[0], q[T.dx], T.fx);
  const float c10 = flerp(q[T.dy], q[T.dy + T.dx], T.fx);
  const float c01 = flerp(q[T.dz], q[T.dz + T.dx], T.fx);
  const float c11 = flerp(q[T.dz + T.dy], q[T.dz + T.dy + T.dx], T.fx);
  T.g = flerp(flerp(c00, c10, T.fy), flerp(c01, c11, T.fy), T.fz);
  return T;
}

struct SampleEval {
  bool contrib;
  float a, b, k, kbase;
  float4 s;
  uint32_t i0, i1;
  float f;
  float cr, cg, cb;  // composited colour
  float m;           // shading scale of s.rgb (1 when unshaded)
  GridTap gt;
  V3 p;
};

// One loop iteration's parameter-dependent evaluation at t (raycast.slang
// 770-1019 without the empty-space skip). `shade_now` computes the 6-tap
// shading only when the sample contributes.
template <bool SHADED, class SamplerK>
__device__ __forceinline__ SampleEval eval_sample(const SamplerK &S, const DiffArgs &A, const float4 *tf,
                                                  const Ray &R, const r3d_frame_params &pc, float t,
                                                  float *lod_step_out) {
  SampleEval E;
  const float foot = R.ortho ? R.cone : t * R.cone;
  const float lod = fclamp_ord(log2_ge1(fmax_sel(foot / R.voxel, 1.0f)) + pc.lod_bias, 0.0f, R.max_lod);
  const float lod_step = fmax_sel(exp2_core(lod), R.span_step);
  *lod_step_out = lod_step;
  const V3 p = vadd(R.ro, vscale(R.dir, t));
  E.p = p;
  const float v = S.sample(p, lod);
  E.contrib = false;
  if (v < pc.threshold) return E;  // low cut: the sample is void
  const float tt = fclamp_ord(v * 256.0f - 0.5f, 0.0f, 255.0f);
  E.i0 = (uint32_t)tt;
  E.i1 = E.i0 + 1u < 256u ? E.i0 + 1u : 255u;
  E.f = tt - (float)E.i0;
  const float4 q0 = tf[E.i0], q1 = tf[E.i1];
  E.s = make_float4(flerp(q0.x, q1.x, E.f), flerp(q0.y, q1.y, E.f), flerp(q0.z, q1.z, E.f), flerp(q0.w, q1.w, E.f));
  E.gt = grid_tap(A.grid, A.gx, A.gy, A.gz, p);
  E.kbase = R.opac_k * lod_step;
  E.k = E.kbase * E.gt.g;
  E.b = 1.0f - fsat(E.s.w);
  E.a = 1.0f - pow_unit(E.b, E.k);
  if (!(E.a > 0.003f)) return E;
  E.contrib = true;
  float cr = E.s.x, cg = E.s.y, cb = E.s.z;
  E.m = 1.0f;
  if (SHADED) {
    const float e = R.voxel * lod_step;
    const float gx = S.sample(mk(p.x + e, p.y, p.z), lod) - S.sample(mk(p.x - e, p.y, p.z), lod);
    const float gy = S.sample(mk(p.x, p.y + e, p.z), lod) - S.sample(mk(p.x, p.y - e, p.z), lod);
    const float gz = S.sample(mk(p.x, p.y, p.z + e), lod) - S.sample(mk(p.x, p.y, p.z - e), lod);
    const V3 g = mk(gx, gy, gz);
    const float gm = vlen(g);
    if (gm > 1e-4f) {
      const V3 key = mk(0x1.a20bd8p-2f, 0x1.a20bd8p-1f, -0x1.a20bd8p-2f);
      const V3 n = mk(g.x / gm, g.y / gm, g.z / gm);
      const float diff = 0.7f * __builtin_fabsf(vdot(n, R.lvec)) + 0.3f * __builtin_fabsf(vdot(n, key));
      const float spec = pow24(__builtin_fabsf(vdot(n, R.hvec))) * 0.3f;
      const float lit = 0.25f + 0.75f * diff;
      const float w = fsat(gm * 6.0f);
      const float m = flerp(1.0f, lit, w);
      const float sw = spec * w;
      E.m = m;
      cr = cr * m + sw;
      cg = cg * m + sw;
      cb = cb * m + sw;
    }
  }
  E.cr = cr;
  E.cg = cg;
  E.cb = cb;
  return E;
}

__device__ __forceinline__ bool tile_pixel(const DiffArgs &A, uint32_t tile, uint32_t lane, uint32_t *v_out,
                                           uint32_t *lx, uint32_t *ly, uint32_t *ray) {
  uint32_t v = 0;
  while (v + 1u < A.nviews && A.tile_start[v + 1] <= tile) v++;
  const r3d_frame_params &pc = A.views[v];
  const bool half = (pc.view_flags & 32u) != 0u;
  const uint32_t W = pc.viewport[0], H = pc.viewport[1];
  const uint32_t TW = half ? (W + 1u) / 2u : W, TH = half ? (H + 1u) / 2u : H;
  const uint32_t tiles_x = (TW + TILE - 1u) / TILE;
  const uint32_t lt = tile - A.tile_start[v];
  const uint32_t tx = (lt % tiles_x) * TILE + (lane % TILE);
  const uint32_t ty = (lt / tiles_x) * TILE + (lane / TILE);
  *v_out = v;
  if (tx >= TW || ty >= TH) return false;
  *lx = half ? tx * 2u : tx;
  *ly = half ? ty * 2u : ty;
  *ray = A.ray_start[v] + ty * TW + tx;
  return true;
}

// REC: write the checkpoints (a forward that an adjoint will follow); without
// them the kernel renders the same image and keeps nothing.
template <bool SHADED, bool REC, bool WIDE>
__global__ __launch_bounds__(64) void k_diff_forward(DiffArgs A) {
  __shared__ float4 s_tf[256];
  __shared__ LevelDev s_lev[R3D_MAX_LEVELS];
  const uint32_t lane = threadIdx.x;
  for (uint32_t i = lane; i < 256u; i += 64u)
    s_tf[i] = make_float4(A.tf[i * 4 + 0], A.tf[i * 4 + 1], A.tf[i * 4 + 2], A.tf[i * 4 + 3]);
  if (lane < A.vol.nlev) s_lev[lane] = A.vol.levels[lane];
  __syncthreads();
  uint32_t v, px, py, ray;
  if (!tile_pixel(A, blockIdx.x, lane, &v, &px, &py, &ray)) return;
  const r3d_frame_params &pc = A.views[v];
  const SamplerT<WIDE> S{A.vol.base, s_lev, A.vol.nlev};
  const Ray R = setup_ray(A.vol, pc, px, py);
  RayRecord *rec = REC ? A.rec + ray : nullptr;
  float ar = 0.0f, ag = 0.0f, ab = 0.0f, aa = 0.0f;
  uint32_t steps = 0;
  if (R.hit) {
    float t = R.t0;
    while (t < R.tfar && steps < MAX_STEPS) {
      if (REC && (steps & (CKPT - 1u)) == 0u) rec->ck[steps / CKPT] = make_float2(t, aa);
      steps++;
      float lod_step;
      const SampleEval E = eval_sample<SHADED>(S, A, s_tf, R, pc, t, &lod_step);
      if (E.contrib) {
        const float wa = (1.0f - aa) * E.a;
        ar = ar + wa * E.cr;
        ag = ag + wa * E.cg;
        ab = ab + wa * E.cb;
        aa = aa + wa;
        if (aa >= A.cutoff) break;
      }
      t += R.dt * lod_step;
    }
  }
  if (REC) rec->nsteps = steps;
  const float ta = 1.0f - aa;
  const float4 c = R.hit ? make_float4(ar + ta * 0.02f, ag + ta * 0.02f, ab + ta * 0.03f, aa)
                         : make_float4(0.02f, 0.02f, 0.03f, 0.0f);
  const bool half = (pc.view_flags & 32u) != 0u;
  const uint32_t ox = pc.view_org & 0xffffu, oy = pc.view_org >> 16;
  const uint32_t rep = half ? 2u : 1u;
  for (uint32_t dy = 0; dy < rep; dy++)
    for (uint32_t dx = 0; dx < rep; dx++) {
      const uint32_t qx = px + dx, qy = py + dy;
      if (qx >= pc.viewport[0] || qy >= pc.viewport[1]) continue;
      const uint32_t gx = ox + qx, gy = oy + qy;
      if (gx >= A.img_w || gy >= A.img_h) continue;
      ((float4 *)A.img)[(size_t)gy * A.img_w + gx] = c;
    }
}

// Grid-gradient accumulator: consecutive samples of a ray usually fall in
// the same grid cell, so the eight corner contributions are summed in
// registers and flushed with atomics only when the cell changes.
struct GridAcc {
  uint32_t i000 = 0xffffffffu, dx = 0, dy = 0, dz = 0;
  float w[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  __device__ __forceinline__ void flush(float *dgrid) {
    if (i000 == 0xffffffffu) return;
    float *d = dgrid + i000;
    atomicAdd(d, w[0]);
    atomicAdd(d + dx, w[1]);
    atomicAdd(d + dy, w[2]);
    atomicAdd(d + dy + dx, w[3]);
    atomicAdd(d + dz, w[4]);
    atomicAdd(d + dz + dx, w[5]);
    atomicAdd(d + dz + dy, w[6]);
    atomicAdd(d + dz + dy + dx, w[7]);
    i000 = 0xffffffffu;
  }
  __device__ __forceinline__ void add(const GridTap &g, float dg, float *dgrid) {
    if (g.i000 != i000) {
      flush(dgrid);
      i000 = g.i000;
      dx = g.dx;
      dy = g.dy;
      dz = g.dz;
#pragma unroll
      for (int c = 0; c < 8; c++) w[c] = 0.0f;
    }
    const float ux = 1.0f - g.fx, uy = 1.0f - g.fy, uz = 1.0f - g.fz;
    w[0] += dg * ux * uy * uz;
    w[1] += dg * g.fx * uy * uz;
    w[2] += dg * ux * g.fy * uz;
    w[3] += dg * g.fx * g.fy * uz;
    w[4] += dg * ux * uy * g.fz;
    w[5] += dg * g.fx * uy * g.fz;
    w[6] += dg * ux * g.fy * g.fz;
    w[7] += dg * g.fx * g.fy * g.fz;
  }
};

// Persistent: each 256-lane workgroup walks groups of four 8x8 tiles (one per
// wave), so the per-block TF partials stay few.
template <bool SHADED, bool WIDE>
__global__ __launch_bounds__(256) void k_adjoint(DiffArgs A) {
  __shared__ float4 s_tf[256];
  __shared__ float s_gtf[1024];
  __shared__ LevelDev s_lev[R3D_MAX_LEVELS];
  const uint32_t tid = threadIdx.x;
  for (uint32_t i = tid; i < 256u; i += 256u)
    s_tf[i] = make_float4(A.tf[i * 4 + 0], A.tf[i * 4 + 1], A.tf[i * 4 + 2], A.tf[i * 4 + 3]);
  for (uint32_t i = tid; i < 1024u; i += 256u) s_gtf[i] = 0.0f;
  if (tid < A.vol.nlev) s_lev[tid] = A.vol.levels[tid];
  __syncthreads();
  const SamplerT<WIDE> S{A.vol.base, s_lev, A.vol.nlev};
  const uint32_t ngroups = (A.ntiles + 3u) / 4u;
  for (uint32_t grp = blockIdx.x; grp < ngroups; grp += gridDim.x) {
    const uint32_t tile = grp * 4u + (tid >> 6);
    uint32_t v = 0, px = 0, py = 0, ray = 0;
    if (!(tile < A.ntiles && tile_pixel(A, tile, tid & 63u, &v, &px, &py, &ray))) continue;
    const r3d_frame_params &pc = A.views[v];
    const Ray R = setup_ray(A.vol, pc, px, py);
    const RayRecord &rec = A.rec[ray];
    const uint32_t N = rec.nsteps;
    if (!R.hit || N == 0u) continue;
    // dL/d out: sum over the pixels this ray wrote
    float gr = 0.0f, gg = 0.0f, gb = 0.0f, ga = 0.0f;
    const bool half = (pc.view_flags & 32u) != 0u;
    const uint32_t ox = pc.view_org & 0xffffu, oy = pc.view_org >> 16;
    const uint32_t rep = half ? 2u : 1u;
    for (uint32_t dy = 0; dy < rep; dy++)
      for (uint32_t dx = 0; dx < rep; dx++) {
        const uint32_t qx = px + dx, qy = py + dy;
        if (qx >= pc.viewport[0] || qy >= pc.viewport[1]) continue;
        const uint32_t gx = ox + qx, gy = oy + qy;
        if (gx >= A.img_w || gy >= A.img_h) continue;
        const float4 d = ((const float4 *)A.dimg)[(size_t)gy * A.img_w + gx];
        gr += d.x;
        gg += d.y;
        gb += d.z;
        ga += d.w;
      }
    float dT = gr * 0.02f + gg * 0.02f + gb * 0.03f - ga;  // out = C + T bg, alpha = 1 - T
    GridAcc acc;
    float seg_t[CKPT], seg_T[CKPT];  // contributing samples of the segment only
    for (int j = (int)((N - 1u) / CKPT); j >= 0; j--) {
      const uint32_t i_begin = (uint32_t)j * CKPT;
      const uint32_t i_end = i_begin + CKPT < N ? i_begin + CKPT : N;
      // forward recompute of the segment from its checkpoint
      const float2 ck = rec.ck[j];
      float t = ck.x, aa = ck.y;
      uint32_t nc = 0;
      for (uint32_t i = i_begin; i < i_end; i++) {
        const float foot = R.ortho ? R.cone : t * R.cone;
        const float lod = fclamp_ord(log2_ge1(fmax_sel(foot / R.voxel, 1.0f)) + pc.lod_bias, 0.0f, R.max_lod);
        const float lod_step = fmax_sel(exp2_core(lod), R.span_step);
        const V3 p = vadd(R.ro, vscale(R.dir, t));
        const float sv = S.sample(p, lod);
        if (!(sv < pc.threshold)) {
          const float tt = fclamp_ord(sv * 256.0f - 0.5f, 0.0f, 255.0f);
          const uint32_t i0 = (uint32_t)tt;
          const uint32_t i1 = i0 + 1u < 256u ? i0 + 1u : 255u;
          const float f = tt - (float)i0;
          const float sa = flerp(s_tf[i0].w, s_tf[i1].w, f);
          const GridTap gt = grid_tap(A.grid, A.gx, A.gy, A.gz, p);
          const float kk = (R.opac_k * lod_step) * gt.g;
          const float a = 1.0f - pow_unit(1.0f - fsat(sa), kk);
          if (a > 0.003f) {
            seg_t[nc] = t;
            seg_T[nc] = 1.0f - aa;
            nc++;
            aa = aa + (1.0f - aa) * a;
          }
        }
        t += R.dt * lod_step;
      }
      // reverse over the segment's contributing samples
      for (int k = (int)nc - 1; k >= 0; k--) {
        float lod_step;
        const SampleEval E = eval_sample<SHADED>(S, A, s_tf, R, pc, seg_t[k], &lod_step);
        const float T = seg_T[k];
        const float gc = gr * E.cr + gg * E.cg + gb * E.cb;
        const float da = gc * T - dT * T;
        dT = dT * (1.0f - E.a) + E.a * gc;
        float dsa = 0.0f;
        if (E.s.w > 0.0f && E.s.w < 1.0f) dsa = da * E.k * (1.0f - E.a) / E.b;
        const float w1 = E.f, w0 = 1.0f - E.f;
        const float ta = T * E.a * E.m;
        const float dsr = gr * ta, dsg = gg * ta, dsb = gb * ta;
        atomicAdd(&s_gtf[E.i0 * 4 + 0], w0 * dsr);
        atomicAdd(&s_gtf[E.i0 * 4 + 1], w0 * dsg);
        atomicAdd(&s_gtf[E.i0 * 4 + 2], w0 * dsb);
        atomicAdd(&s_gtf[E.i0 * 4 + 3], w0 * dsa);
        atomicAdd(&s_gtf[E.i1 * 4 + 0], w1 * dsr);
        atomicAdd(&s_gtf[E.i1 * 4 + 1], w1 * dsg);
        atomicAdd(&s_gtf[E.i1 * 4 + 2], w1 * dsb);
        atomicAdd(&s_gtf[E.i1 * 4 + 3], w1 * dsa);
        if (E.b > 0.0f && E.b < 1.0f) {
          const float dk = da * (-(1.0f - E.a) * (log2_pinned(E.b) * 0x1.62e430p-1f));
          acc.add(E.gt, dk * E.kbase, A.dgrid);
        }
      }
    }
    acc.flush(A.dgrid);
  }
  __syncthreads();
  // partials are entry-major, [1024][gridDim.x], so the reduction reads rows
  for (uint32_t i = tid; i < 1024u; i += 256u) A.gtf_part[(size_t)i * gridDim.x + blockIdx.x] = s_gtf[i];
}

// dL/dtf[i] = sum over blocks of partial[i][b]: one workgroup per entry, a
// fixed per-lane order and a fixed tree, so the result is deterministic.
__global__ __launch_bounds__(256) void k_reduce_tf(const float *part, uint32_t nblocks, float *dtf) {
  __shared__ float red[256];
  const uint32_t i = blockIdx.x;
  float s = 0.0f;
  for (uint32_t b = threadIdx.x; b < nblocks; b += 256u) s += part[(size_t)i * nblocks + b];
  red[threadIdx.x] = s;
  __syncthreads();
  for (uint32_t w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) dtf[i] = red[0];
}

}  // namespace r3d

using namespace r3d;

namespace {

struct Plan {
  uint32_t nviews = 0, ntiles = 0, nrays = 0;
  bool shaded = false;
};

// need_ws: the adjoint reads the checkpoints; a forward may go without them.
int validate_diff(r3d_hip_ctx *ctx, const r3d_hip_volume *vol, const r3d_frame_params *views, uint32_t nviews,
                  const r3d_diff_params *dp, Plan *plan, bool need_ws) {
  if (!vol || vol->ctx != ctx) return r3d_fail(ctx, R3D_ERR_INVALID, "diff: bad volume");
  if (!views || nviews == 0 || nviews > R3D_MAX_VIEWS_PER_CALL)
    return r3d_fail(ctx, R3D_ERR_INVALID, "diff: nviews %u outside 1..%u", nviews, R3D_MAX_VIEWS_PER_CALL);
  if (!dp || !dp->tf || !dp->grid || (need_ws && !dp->workspace))
    return r3d_fail(ctx, R3D_ERR_INVALID, need_ws ? "adjoint: tf, grid and workspace are required"
                                                  : "diff: tf and grid are required");
  if (dp->grid_dims[0] == 0 || dp->grid_dims[1] == 0 || dp->grid_dims[2] == 0 || dp->grid_dims[0] > 4096 ||
      dp->grid_dims[1] > 4096 || dp->grid_dims[2] > 4096 ||
      (uint64_t)dp->grid_dims[0] * dp->grid_dims[1] * dp->grid_dims[2] > (1ull << 31))
    return r3d_fail(ctx, R3D_ERR_RANGE, "diff: grid dims");
  const uint32_t mode = views[0].mode;
  if (mode != R3D_MODE_FULL && mode != R3D_MODE_FLAT)
    return r3d_fail(ctx, R3D_ERR_UNSUPPORTED, "diff: only FULL and FLAT views are differentiable");
  uint64_t tiles = 0, rays = 0;
  for (uint32_t i = 0; i < nviews; i++) {
    const r3d_frame_params &p = views[i];
    if (p.mode != mode) return r3d_fail(ctx, R3D_ERR_INVALID, "diff: all views of a call share one mode");
    if (p.slab_grid || p.clip_valid || p.brick_mode || (p.view_flags & 2u))
      return r3d_fail(ctx, R3D_ERR_UNSUPPORTED, "diff: view %u is not a cube-mode view", i);
    if (p.viewport[0] == 0 || p.viewport[1] == 0 || p.viewport[0] > 65535u || p.viewport[1] > 65535u)
      return r3d_fail(ctx, R3D_ERR_INVALID, "diff: view %u viewport", i);
    if (!(p.step_voxels > 0.0f) || !(p.max_mip >= 0.0f))
      return r3d_fail(ctx, R3D_ERR_INVALID, "diff: view %u step_voxels / max_mip", i);
    const bool half = (p.view_flags & 32u) != 0u;
    const uint64_t TW = half ? (p.viewport[0] + 1u) / 2u : p.viewport[0];
    const uint64_t TH = half ? (p.viewport[1] + 1u) / 2u : p.viewport[1];
    tiles += ((TW + TILE - 1u) / TILE) * ((TH + TILE - 1u) / TILE);
    rays += TW * TH;
  }
  if (tiles > 0xffffffffull || rays > 0xffffffffull) return r3d_fail(ctx, R3D_ERR_RANGE, "diff: too many rays");
  if (dp->workspace && dp->workspace_bytes < rays * sizeof(RayRecord))
    return r3d_fail(ctx, R3D_ERR_INVALID, "diff: workspace %llu bytes < %llu needed",
                    (unsigned long long)dp->workspace_bytes, (unsigned long long)(rays * sizeof(RayRecord)));
  plan->nviews = nviews;
  plan->ntiles = (uint32_t)tiles;
  plan->nrays = (uint32_t)rays;
  plan->shaded = mode == R3D_MODE_FULL;
  return R3D_OK;
}

int stage_views(r3d_hip_ctx *ctx, const r3d_frame_params *views, uint32_t nviews, hipStream_t s, ViewSlot **slot,
                DiffArgs *A) {
  // [views][tile_start (n+1)][ray_start (n+1)]
  const size_t pb = sizeof(r3d_frame_params) * nviews, tb = sizeof(uint32_t) * (nviews + 1);
  uint8_t buf[sizeof(r3d_frame_params) * R3D_MAX_VIEWS_PER_CALL + 2 * sizeof(uint32_t) * (R3D_MAX_VIEWS_PER_CALL + 1)];
  memcpy(buf, views, pb);
  uint32_t *ts = (uint32_t *)(buf + pb), *rs = (uint32_t *)(buf + pb + tb);
  ts[0] = rs[0] = 0;
  for (uint32_t i = 0; i < nviews; i++) {
    const bool half = (views[i].view_flags & 32u) != 0u;
    const uint32_t TW = half ? (views[i].viewport[0] + 1u) / 2u : views[i].viewport[0];
    const uint32_t TH = half ? (views[i].viewport[1] + 1u) / 2u : views[i].viewport[1];
    ts[i + 1] = ts[i] + ((TW + TILE - 1u) / TILE) * ((TH + TILE - 1u) / TILE);
    rs[i + 1] = rs[i] + TW * TH;
  }
  int rc = stage(ctx, buf, pb + 2 * tb, s, slot);
  if (rc != R3D_OK) return rc;
  A->views = (const r3d_frame_params *)(*slot)->dev;
  A->tile_start = (const uint32_t *)((const uint8_t *)(*slot)->dev + pb);
  A->ray_start = (const uint32_t *)((const uint8_t *)(*slot)->dev + pb + tb);
  return R3D_OK;
}

}  // namespace

extern "C" size_t r3d_hip_diff_workspace_bytes(const r3d_frame_params *views, uint32_t nviews) {
  if (!views) return 0;
  uint64_t rays = 0;
  for (uint32_t i = 0; i < nviews; i++) {
    const bool half = (views[i].view_flags & 32u) != 0u;
    const uint64_t TW = half ? (views[i].viewport[0] + 1u) / 2u : views[i].viewport[0];
    const uint64_t TH = half ? (views[i].viewport[1] + 1u) / 2u : views[i].viewport[1];
    rays += TW * TH;
  }
  return (size_t)(rays * sizeof(RayRecord));
}

extern "C" int r3d_hip_raycast_diff(r3d_hip_ctx *ctx, const r3d_hip_volume *vol, const r3d_frame_params *views,
                                    uint32_t nviews, const r3d_hip_image *out, const r3d_diff_params *dp,
                                    void *stream) {
  if (!ctx) return R3D_ERR_INVALID;
  Plan plan;
  int rc = validate_diff(ctx, vol, views, nviews, dp, &plan, false);
  if (rc != R3D_OK) return rc;
  r3d_panes_reset(ctx);
  if (!out || !out->rgba || out->width == 0 || out->height == 0)
    return r3d_fail(ctx, R3D_ERR_INVALID, "diff: FP32 output image required");
  R3D_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  hipStream_t s = (hipStream_t)stream;
  DiffArgs A;
  memset(&A, 0, sizeof A);
  ViewSlot *slot = nullptr;
  rc = stage_views(ctx, views, nviews, s, &slot, &A);
  if (rc != R3D_OK) return rc;
  A.vol = vol_dev(vol);
  A.nviews = nviews;
  A.ntiles = plan.ntiles;
  A.tf = dp->tf;
  A.grid = dp->grid;
  A.gx = dp->grid_dims[0];
  A.gy = dp->grid_dims[1];
  A.gz = dp->grid_dims[2];
  A.cutoff = dp->alpha_cutoff > 0.0f ? dp->alpha_cutoff : 0.98f;
  A.img = out->rgba;
  A.img_w = out->width;
  A.img_h = out->height;
  A.rec = (RayRecord *)dp->workspace;
  typedef void (*FwdKernel)(DiffArgs);
  // [64-bit offsets][shaded][checkpoints]: 32-bit offsets for a pyramid below 4 GiB
  static const FwdKernel fwd[2][2][2] = {
      {{k_diff_forward<false, false, false>, k_diff_forward<false, true, false>},
       {k_diff_forward<true, false, false>, k_diff_forward<true, true, false>}},
      {{k_diff_forward<false, false, true>, k_diff_forward<false, true, true>},
       {k_diff_forward<true, false, true>, k_diff_forward<true, true, true>}}};
  hipLaunchKernelGGL(fwd[A.vol.narrow ? 0 : 1][plan.shaded ? 1 : 0][A.rec ? 1 : 0], dim3(plan.ntiles), dim3(64), 0,
                     s, A);
  R3D_HIP_CHECK(ctx, hipGetLastError());
  R3D_HIP_CHECK(ctx, hipEventRecord(slot->ev, s));
  slot->recorded = true;
  return R3D_OK;
}

extern "C" int r3d_hip_raycast_adjoint(r3d_hip_ctx *ctx, const r3d_hip_volume *vol, const r3d_frame_params *views,
                                       uint32_t nviews, const float *dL_drgba, uint32_t img_w, uint32_t img_h,
                                       float *dL_dtf, float *dL_dgrid, const r3d_diff_params *dp, void *stream) {
  if (!ctx) return R3D_ERR_INVALID;
  Plan plan;
  int rc = validate_diff(ctx, vol, views, nviews, dp, &plan, true);
  if (rc != R3D_OK) return rc;
  if (!dL_drgba || !dL_dtf || !dL_dgrid || img_w == 0 || img_h == 0)
    return r3d_fail(ctx, R3D_ERR_INVALID, "adjoint: dL_drgba, dL_dtf, dL_dgrid and the image size are required");
  R3D_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  hipStream_t s = (hipStream_t)stream;
  const uint32_t ngroups = (plan.ntiles + 3u) / 4u;
  const uint32_t cap = (uint32_t)(ctx->num_cus > 0 ? ctx->num_cus : 256) * 8u;
  const uint32_t nblocks = ngroups < cap ? ngroups : cap;
  const size_t need = (size_t)nblocks * 1024u * sizeof(float);
  if (ctx->gtf_cap < need) {
    if (ctx->d_gtf) {
      R3D_HIP_CHECK(ctx, hipStreamSynchronize(s));
      R3D_HIP_CHECK(ctx, hipFree(ctx->d_gtf));
      ctx->d_gtf = nullptr;
      ctx->gtf_cap = 0;
    }
    if (hipMalloc(&ctx->d_gtf, need) != hipSuccess)
      return r3d_fail(ctx, R3D_ERR_NOMEM, "adjoint: TF partials %zu bytes", need);
    ctx->gtf_cap = need;
  }
  DiffArgs A;
  memset(&A, 0, sizeof A);
  ViewSlot *slot = nullptr;
  rc = stage_views(ctx, views, nviews, s, &slot, &A);
  if (rc != R3D_OK) return rc;
  A.vol = vol_dev(vol);
  A.nviews = nviews;
  A.ntiles = plan.ntiles;
  A.tf = dp->tf;
  A.grid = dp->grid;
  A.gx = dp->grid_dims[0];
  A.gy = dp->grid_dims[1];
  A.gz = dp->grid_dims[2];
  A.cutoff = dp->alpha_cutoff > 0.0f ? dp->alpha_cutoff : 0.98f;
  A.dimg = dL_drgba;
  A.img_w = img_w;
  A.img_h = img_h;
  A.rec = (RayRecord *)dp->workspace;
  A.gtf_part = ctx->d_gtf;
  A.dgrid = dL_dgrid;
  typedef void (*AdjKernel)(DiffArgs);
  static const AdjKernel adj[2][2] = {{k_adjoint<false, false>, k_adjoint<true, false>},
                                      {k_adjoint<false, true>, k_adjoint<true, true>}};
  hipLaunchKernelGGL(adj[A.vol.narrow ? 0 : 1][plan.shaded ? 1 : 0], dim3(nblocks), dim3(256), 0, s, A);
  hipLaunchKernelGGL(k_reduce_tf, dim3(1024), dim3(256), 0, s, (const float *)ctx->d_gtf, nblocks, dL_dtf);
  R3D_HIP_CHECK(ctx, hipGetLastError());
  R3D_HIP_CHECK(ctx, hipEventRecord(slot->ev, s));
  slot->recorded = true;
  return R3D_OK;
}
