// r3d_deblock.hip -- 3ddct's decode-side deblocking post-filter on gfx950
// (tools/3ddct/deblock.c), u8.
//
// Volume form: the reference filters every internal 16-aligned plane in
// place, x seams, then y seams, then z seams. Within one pass the seams are 16
// voxels apart and each touches only the two layers on either side, so every
// seam sample of a pass is independent. Rows that are 16-byte aligned take
// one fused kernel over tiles that each hold a whole seam neighbourhood per
// axis (the volume crosses HBM once); any other alignment takes one kernel per
// pass, one lane per sample (the passes stay in order). Chunk form: one
// workgroup per chunk filters its six faces in order against pristine
// neighbours, with the chunk's running values in LDS. The arithmetic is the
// reference's FP32 sequence (unfused, correctly rounded division, roundf,
// clamp), so the output is bit-identical. HBM-bound byte work.
#include <type_traits>

#include "r3d_dmath.h"
#include "r3d_internal.h"

namespace r3d {

constexpr uint32_t DB_N = 16u;

// deblock.c:30-45 without the two activity tests: the nudges (d0, d1) of
// a p0|q0 step d = q0 - p0, (0, 0) for a real edge or no step. d0 is never 0
// otherwise (w > 0).
__device__ __forceinline__ float2 db_step_nudge(float d, float alpha) {
  const float ad = __builtin_fabsf(d);
  if (ad >= alpha || ad == 0.0f) return make_float2(0.0f, 0.0f);
  float d0 = 0.40f * d;
  const float w = 1.0f - ad / alpha;
  d0 *= w;
  return make_float2(d0, 0.4f * d0);
}

// deblock.c:26-46: the nudge of the p0|q0 step, 0 for a real edge. The three
// tests have no side effects, so any order of them gives the same result.
__device__ __forceinline__ float db_delta(float p1, float p0, float q0, float q1, float alpha, float beta, float *d1) {
  *d1 = 0.0f;
  if (__builtin_fabsf(p1 - p0) >= beta || __builtin_fabsf(q1 - q0) >= beta) return 0.0f;
  const float2 n = db_step_nudge(q0 - p0, alpha);
  *d1 = n.y;
  return n.x;
}

// Voxels are u8, so the step q0 - p0 is one of the 511 integers -255 .. 255
// and the nudges are a table of it: entry d + 255, computed with the
// reference's own sequence (correctly rounded division included) once per
// workgroup. One LDS read replaces the division and three products per sample.
constexpr uint32_t DB_NSTEP = 511u;
__device__ __forceinline__ void db_fill_nudges(float2 *tab, uint32_t t, uint32_t nt, float alpha) {
  for (uint32_t i = t; i < DB_NSTEP; i += nt) tab[i] = db_step_nudge((float)((int32_t)i - 255), alpha);
}

__device__ __forceinline__ uint8_t db_u8(float v) {
  return (uint8_t)roundf(v < 0.0f ? 0.0f : v > 255.0f ? 255.0f : v);
}

// deblock.c:48-67
__host__ __device__ __forceinline__ void db_limits(float step, float range, float strength, float *alpha,
                                                   float *beta) {
  if (step > 0.0f) {
    *alpha = 2.0f * step * strength;
    *beta = 1.2f * step * strength;
  } else {
    *alpha = 0.06f * range * strength;
    *beta = 0.04f * range * strength;
  }
}

struct DbLim {
  float alpha, beta;
  uint32_t on;
};

// deblock.c:144-152: the data range over the samples i = 0, s, 2s, ... with
// s = max(1, total / 4096), then the fallback thresholds.
__global__ __launch_bounds__(1024) void k_db_range(const uint8_t *vol, uint64_t total, float strength, DbLim *lim) {
  __shared__ uint32_t s_mn[16], s_mx[16];
  const uint64_t stride = total / 4096u ? total / 4096u : 1u;
  const uint64_t ns = (total + stride - 1u) / stride;
  uint32_t mn = 255u, mx = 0u;
  for (uint64_t k = threadIdx.x; k < ns; k += 1024u) {
    const uint32_t v = vol[k * stride];
    mn = v < mn ? v : mn;
    mx = v > mx ? v : mx;
  }
  for (int o = 32; o > 0; o >>= 1) {
    const uint32_t a = __shfl_xor(mn, o), b = __shfl_xor(mx, o);
    mn = a < mn ? a : mn;
    mx = b > mx ? b : mx;
  }
  if ((threadIdx.x & 63u) == 0u) {
    s_mn[threadIdx.x >> 6] = mn;
    s_mx[threadIdx.x >> 6] = mx;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < 16; w++) {
      mn = s_mn[w] < mn ? s_mn[w] : mn;
      mx = s_mx[w] > mx ? s_mx[w] : mx;
    }
    const float range = (float)mx - (float)mn;
    DbLim l;
    l.on = range > 0.0f ? 1u : 0u;
    db_limits(0.0f, range, strength, &l.alpha, &l.beta);
    *lim = l;
  }
}

struct DbPass {
  uint8_t *vol;
  uint64_t nz, ny, nx;
  uint64_t nseam;   // seams along the pass's axis (planes 16, 32, ... with plane + 1 < extent)
  uint64_t total;   // work items of the pass (seam samples, or groups of W of them)
  DbLim lim;        // used when flim is null
  const DbLim *flim;
};

// W consecutive seam samples along x: the four layers p1, p0 | q0, q1 as W-byte
// vectors (v points at q0, s is the stride across the seam). A layer is
// stored back only if one of its bytes changed, and no other lane of the pass
// touches its bytes.
template <int W>
__device__ __forceinline__ void db_filter_vec(uint8_t *v, int64_t s, float alpha, float beta) {
  using V = typename std::conditional<W == 16, uint4, typename std::conditional<W == 4, uint32_t, uint8_t>::type>::type;
  V r[4];
  r[0] = *(const V *)(v - 2 * s);
  r[1] = *(const V *)(v - s);
  r[2] = *(const V *)v;
  r[3] = *(const V *)(v + s);
  uint8_t b[4][W];
#pragma unroll
  for (int k = 0; k < 4; k++) __builtin_memcpy(b[k], &r[k], W);
  bool ch[4] = {false, false, false, false};
#pragma unroll
  for (int x = 0; x < W; x++) {
    const float p1 = b[0][x], p0 = b[1][x], q0 = b[2][x], q1 = b[3][x];
    float d1;
    const float d0 = db_delta(p1, p0, q0, q1, alpha, beta, &d1);
    if (d0 != 0.0f) {
      b[1][x] = db_u8(p0 + d0);
      b[2][x] = db_u8(q0 - d0);
      b[0][x] = db_u8(p1 + d1);
      b[3][x] = db_u8(q1 - d1);
      ch[0] = ch[1] = ch[2] = ch[3] = true;
    }
  }
  if (ch[0]) {
#pragma unroll
    for (int k = 0; k < 4; k++) __builtin_memcpy(&r[k], b[k], W);
    *(V *)(v - 2 * s) = r[0];
    *(V *)(v - s) = r[1];
    *(V *)v = r[2];
    *(V *)(v + s) = r[3];
  }
}

// x seams, rows 4-byte aligned: bytes x-4..x-1 and x..x+3 of the row as two
// aligned dwords (x = 16 (j + 1)).
__device__ __forceinline__ void db_filter_x4(uint8_t *q, float alpha, float beta) {
  const uint32_t lo = *(const uint32_t *)(q - 4), hi = *(const uint32_t *)q;
  const float p1 = (float)((lo >> 16) & 255u), p0 = (float)(lo >> 24), q0 = (float)(hi & 255u),
              q1 = (float)((hi >> 8) & 255u);
  float d1;
  const float d0 = db_delta(p1, p0, q0, q1, alpha, beta, &d1);
  if (d0 == 0.0f) return;
  const uint32_t np1 = db_u8(p1 + d1), np0 = db_u8(p0 + d0), nq0 = db_u8(q0 - d0), nq1 = db_u8(q1 - d1);
  *(uint32_t *)(q - 4) = (lo & 0x0000FFFFu) | (np1 << 16) | (np0 << 24);
  *(uint32_t *)q = (hi & 0xFFFF0000u) | nq0 | (nq1 << 8);
}

// AXIS 0: x seams, item i = (row, j); W = 4 uses aligned dwords, W = 1 bytes.
// AXIS 1: y seams, item i = ((z, j), x / W). AXIS 2: z seams, i = ((j, y), x / W).
// I is the index type: 32-bit divisions when the pass has fewer than 2^32
// items (a 64-bit division by a runtime value costs far more than the filter).
template <int AXIS, int W, typename I>
__device__ __forceinline__ void db_item(const DbPass &a, I i, const DbLim &l) {
  const I nseam = (I)a.nseam, nx = (I)a.nx, ny = (I)a.ny;
  if (AXIS == 0) {
    const I row = i / nseam, j = i - row * nseam;
    uint8_t *q = a.vol + (uint64_t)row * a.nx + (uint64_t)(j + 1u) * DB_N;
    if (W == 4)
      db_filter_x4(q, l.alpha, l.beta);
    else
      db_filter_vec<1>(q, 1, l.alpha, l.beta);
  } else {
    const I nxw = nx / (I)W;
    const I r = i / nxw, xw = i - r * nxw;
    if (AXIS == 1) {
      const I z = r / nseam, j = r - z * nseam;
      db_filter_vec<W>(a.vol + ((uint64_t)z * a.ny + (uint64_t)(j + 1u) * DB_N) * a.nx + (uint64_t)xw * W,
                       (int64_t)a.nx, l.alpha, l.beta);
    } else {
      const I j = r / ny, y = r - j * ny;
      db_filter_vec<W>(a.vol + ((uint64_t)(j + 1u) * DB_N * a.ny + y) * a.nx + (uint64_t)xw * W,
                       (int64_t)(a.ny * a.nx), l.alpha, l.beta);
    }
  }
}

template <int AXIS, int W>
__global__ __launch_bounds__(256) void k_db_pass(DbPass a) {
  const DbLim l = a.flim ? *a.flim : a.lim;
  if (!l.on) return;
  if (a.total <= 0xFFFFFFFFull && a.nx <= 0xFFFFFFFFull && a.ny <= 0xFFFFFFFFull) {
    const uint32_t step = gridDim.x * 256u;
    for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < (uint32_t)a.total; i += step) {
      db_item<AXIS, W, uint32_t>(a, i, l);
      if (i > 0xFFFFFFFFu - step) break;  // the next index would wrap
    }
  } else {
    const uint64_t step = (uint64_t)gridDim.x * 256u;
    for (uint64_t i = (uint64_t)blockIdx.x * 256u + threadIdx.x; i < a.total; i += step) db_item<AXIS, W, uint64_t>(a, i, l);
  }
}

// All three passes in one kernel for rows that are 16-byte aligned, over
// tiles of 128 x 16 x 16 voxels shifted by -2 on every axis (x range
// [128m - 2, 128m + 126), y [16j - 2, 16j + 14), z likewise): a tile holds the
// whole two-layer neighbourhood of every seam it covers (8 along x, one along
// y and z), every voxel belongs to one tile, and a tile reads no byte that
// another tile writes (tiles run concurrently, in any order). Eight lanes
// share a 128-byte row, 16 bytes each (coalesced loads); lane c of a row
// holds tile bytes 16c .. 16c + 15, whose first four are the x seam
// 128m + 16c, so the x seams are filtered in registers (the two bytes before
// each lane's load come from its neighbour lane). The y and z seams follow
// through the tile in LDS, a dword of four columns per lane, in the
// reference's order, and only changed dwords go back. Every sample takes its
// nudges from the workgroup's table of the 511 possible steps and writes its
// bytes with round_u8_into. The volume is read once; the per-pass kernels
// read it whole for the x pass (seams every 16 bytes touch every line) and a
// quarter again for each of y and z.
struct DbTile {
  uint8_t *vol;
  uint32_t nz, ny, nx;
  uint32_t tx, ty, tz;  // tiles per axis
  DbLim lim;
  const DbLim *flim;
};

// One seam sample from its four layers (u8 values) and the nudge table:
// whether it is filtered, and its nudges.
__device__ __forceinline__ bool db_nudge(uint32_t b0, uint32_t b1, uint32_t b2, uint32_t b3, float beta,
                                         const float2 *tab, float2 *n) {
  *n = tab[b2 + 255u - b1];
  const float p1 = (float)b0, p0 = (float)b1, q0 = (float)b2, q1 = (float)b3;
  return n->x != 0.0f && __builtin_fabsf(p1 - p0) < beta && __builtin_fabsf(q1 - q0) < beta;
}

// An x seam sample held in one dword, bytes p1 | p0 | q0 | q1.
__device__ __forceinline__ uint32_t db_sample_dword(uint32_t r, float beta, const float2 *tab) {
  const uint32_t b0 = r & 255u, b1 = (r >> 8) & 255u, b2 = (r >> 16) & 255u, b3 = r >> 24;
  float2 n;
  if (!db_nudge(b0, b1, b2, b3, beta, tab, &n)) return r;
  uint32_t o = round_u8_into((float)b0 + n.y, 0u, 0u);
  o = round_u8_into((float)b1 + n.x, 1u, o);
  o = round_u8_into((float)b2 - n.x, 2u, o);
  return round_u8_into((float)b3 - n.y, 3u, o);
}

// Four columns packed in one dword per layer (p1 | p0 | q0 | q1 across the
// seam); byte k is filtered when bit k of `valid` is set.
__device__ __forceinline__ void db_sample4(uint32_t &w0, uint32_t &w1, uint32_t &w2, uint32_t &w3, uint32_t valid,
                                           float beta, const float2 *tab) {
  uint32_t o0 = w0, o1 = w1, o2 = w2, o3 = w3;
#pragma unroll
  for (uint32_t k = 0; k < 4u; k++) {
    const uint32_t b0 = (w0 >> (8u * k)) & 255u, b1 = (w1 >> (8u * k)) & 255u, b2 = (w2 >> (8u * k)) & 255u,
                   b3 = (w3 >> (8u * k)) & 255u;
    float2 n;
    if (((valid >> k) & 1u) && db_nudge(b0, b1, b2, b3, beta, tab, &n)) {
      o0 = round_u8_into((float)b0 + n.y, k, o0);
      o1 = round_u8_into((float)b1 + n.x, k, o1);
      o2 = round_u8_into((float)b2 - n.x, k, o2);
      o3 = round_u8_into((float)b3 - n.y, k, o3);
    }
  }
  w0 = o0;
  w1 = o1;
  w2 = o2;
  w3 = o3;
}

// Tile width in x. 64-wide tiles (20 KiB of LDS, 8 workgroups per CU) read
// faster but write slower, two workgroups dirtying each 128-byte line:
// 0.58 ms against 0.49 on a 1024^3 region with 102 M changed voxels.
constexpr uint32_t DBT_W = 128u;

// Tile width in x: TW / 16 lanes share a row, each row of the tile's 256 (16 y
// x 16 z) is held by one lane group.
template <uint32_t TW>
__global__ __launch_bounds__(256) void k_db_tiles(DbTile a) {
  constexpr uint32_t LPR = TW / 16u;    // lanes per row
  constexpr uint32_t RPI = 256u / LPR;  // rows per pass of the workgroup
  constexpr uint32_t NQ = 256u / RPI;   // rows per lane
  constexpr uint32_t RW = TW / 4u;      // dwords per row
  constexpr uint32_t NC = RW * 16u / 256u;  // y or z seam dword columns per lane
  static_assert(TW % 64u == 0u && TW <= 256u, "whole rows per lane group and seam columns per lane");
  __shared__ __attribute__((aligned(16))) uint32_t s_t[16 * 16 * RW];  // [lz][ly][lx / 4]
  __shared__ float2 s_nudge[DB_NSTEP];
  const DbLim l = a.flim ? *a.flim : a.lim;
  if (!l.on) return;
  // XCD bands: a row's x tiles run on one XCD one after another, so the line
  // holding a tile's first two bytes (also the tile before it's) is an L2 hit;
  // in blockIdx order it was fetched again, doubling the reads
  const uint32_t b = xcd_band(blockIdx.x, gridDim.x);
  const uint32_t tm = b % a.tx, tj = (b / a.tx) % a.ty, tk = b / (a.tx * a.ty);
  const uint32_t t = threadIdx.x, c = t % LPR, rs = t / LPR;  // 16-byte column, first row
  const uint32_t xb = TW * tm;  // the tile starts at x = xb - 2
  const int64_t y0 = 16 * (int64_t)tj - 2, z0 = 16 * (int64_t)tk - 2;
  const uint32_t xc = xb + 16u * c;  // x of this lane's load
  const bool col_in = xc < a.nx, has_lo = c == 0u && tm > 0u;
  // rows r = rs + RPI q of the tile (r = ly + 16 lz)
  uint4 W[NQ];
  uint32_t lo[NQ];
  bool rin[NQ];
  uint8_t *rowp[NQ];
#pragma unroll
  for (uint32_t q = 0; q < NQ; q++) {
    const uint32_t r = rs + RPI * q;
    const int64_t gy = y0 + (r & 15u), gz = z0 + (r >> 4);
    rin[q] = gy >= 0 && gy < a.ny && gz >= 0 && gz < a.nz;
    rowp[q] = a.vol + ((uint64_t)(rin[q] ? gz : 0) * a.ny + (uint64_t)(rin[q] ? gy : 0)) * a.nx;
    W[q] = make_uint4(0, 0, 0, 0);
    lo[q] = 0;
    if (rin[q] && col_in) W[q] = *(const uint4 *)(rowp[q] + xc);
    if (rin[q] && has_lo) lo[q] = *(const uint32_t *)(rowp[q] + xb - 4u);
  }
  db_fill_nudges(s_nudge, t, 256u, l.alpha);
  __syncthreads();
  // this lane's x seam is xc (xc + 1 < nx follows from col_in: nx is a multiple of 16)
  const bool seam_x = col_in && xc >= 16u;
#pragma unroll
  for (uint32_t q = 0; q < NQ; q++) {
    const uint32_t up = __shfl_up(W[q].w, 1u, LPR);
    const uint32_t pw = c == 0u ? lo[q] : up;
    uint32_t r0 = (pw >> 16) | (W[q].x << 16);
    const uint32_t r1 = (W[q].x >> 16) | (W[q].y << 16), r2 = (W[q].y >> 16) | (W[q].z << 16),
                   r3 = (W[q].z >> 16) | (W[q].w << 16);
    if (rin[q] && seam_x) r0 = db_sample_dword(r0, l.beta, s_nudge);
    const uint32_t r = rs + RPI * q;
    *(uint4 *)(s_t + r * RW + 4u * c) = make_uint4(r0, r1, r2, r3);
  }
  __syncthreads();
  // y seam at 16j: dword columns (x4, lz)
  if (tj > 0u && 16u * tj + 1u < a.ny) {
#pragma unroll
    for (uint32_t q = 0; q < NC; q++) {
      const uint32_t it = t + 256u * q, x4 = it % RW, cz = it / RW;
      const int64_t gzc = z0 + cz;
      if (gzc >= 0 && gzc < a.nz) {
        uint32_t valid = 0;
#pragma unroll
        for (int k = 0; k < 4; k++) {
          const int64_t gx = (int64_t)xb - 2 + 4 * x4 + k;
          valid |= (gx >= 0 && gx < a.nx ? 1u : 0u) << k;
        }
        uint32_t *cc = s_t + cz * 16u * RW + x4;
        uint32_t w0 = cc[0], w1 = cc[RW], w2 = cc[2u * RW], w3 = cc[3u * RW];
        db_sample4(w0, w1, w2, w3, valid, l.beta, s_nudge);
        cc[0] = w0;
        cc[RW] = w1;
        cc[2u * RW] = w2;
        cc[3u * RW] = w3;
      }
    }
  }
  __syncthreads();
  // z seam at 16k: dword columns (x4, ly)
  if (tk > 0u && 16u * tk + 1u < a.nz) {
#pragma unroll
    for (uint32_t q = 0; q < NC; q++) {
      const uint32_t it = t + 256u * q, x4 = it % RW, cy = it / RW;
      const int64_t gyc = y0 + cy;
      if (gyc >= 0 && gyc < a.ny) {
        uint32_t valid = 0;
#pragma unroll
        for (int k = 0; k < 4; k++) {
          const int64_t gx = (int64_t)xb - 2 + 4 * x4 + k;
          valid |= (gx >= 0 && gx < a.nx ? 1u : 0u) << k;
        }
        uint32_t *cc = s_t + cy * RW + x4;
        constexpr uint32_t PL = 16u * RW;
        uint32_t w0 = cc[0], w1 = cc[PL], w2 = cc[2u * PL], w3 = cc[3u * PL];
        db_sample4(w0, w1, w2, w3, valid, l.beta, s_nudge);
        cc[0] = w0;
        cc[PL] = w1;
        cc[2u * PL] = w2;
        cc[3u * PL] = w3;
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (uint32_t q = 0; q < NQ; q++) {
    const uint32_t r = rs + RPI * q;
    const uint4 f = *(const uint4 *)(s_t + r * RW + 4u * c);
    const uint32_t nf = __shfl_down(f.x, 1u, LPR);  // tile bytes 16c + 16, 16c + 17 (c < LPR - 1)
    if (!rin[q]) continue;
    uint8_t *row = rowp[q];
    // x = xb - 2, xb - 1: tile bytes 0, 1 (the previous tile owns xb - 4, xb - 3)
    if (has_lo && (f.x & 0xFFFFu) != (lo[q] >> 16)) *(uint16_t *)(row + xb - 2u) = (uint16_t)(f.x & 0xFFFFu);
    if (!col_in) continue;
    // x = xc .. xc + 15: tile bytes 16c + 2 .. 16c + 17; for the last column
    // the last two (x = xb + TW - 2, TW - 1) belong to the next tile
    const uint32_t n0 = (f.x >> 16) | (f.y << 16), n1 = (f.y >> 16) | (f.z << 16), n2 = (f.z >> 16) | (f.w << 16),
                   n3 = (f.w >> 16) | (nf << 16);
    uint32_t *d = (uint32_t *)(row + xc);
    if (c + 1u < LPR) {  // all 16 bytes are this tile's: one store when any changed
      if (n0 != W[q].x || n1 != W[q].y || n2 != W[q].z || n3 != W[q].w) *(uint4 *)d = make_uint4(n0, n1, n2, n3);
    } else {
      if (n0 != W[q].x) d[0] = n0;
      if (n1 != W[q].y) d[1] = n1;
      if (n2 != W[q].z) d[2] = n2;
      if ((n3 & 0xFFFFu) != (W[q].w & 0xFFFFu)) *(uint16_t *)(row + xc + 12u) = (uint16_t)(n3 & 0xFFFFu);
    }
  }
}

__device__ __forceinline__ void block_minmax(const uint8_t *c, uint32_t t, uint32_t *s_red, float *mn, float *mx) {
  uint32_t lo = 255u, hi = 0u;
  const uint4 q = ((const uint4 *)c)[t];
  const uint32_t w[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
  for (int k = 0; k < 16; k++) {
    const uint32_t v = (w[k >> 2] >> ((k & 3) * 8)) & 255u;
    lo = v < lo ? v : lo;
    hi = v > hi ? v : hi;
  }
  for (int o = 32; o > 0; o >>= 1) {
    const uint32_t a = __shfl_xor(lo, o), b = __shfl_xor(hi, o);
    lo = a < lo ? a : lo;
    hi = b > hi ? b : hi;
  }
  if ((t & 63u) == 0u) s_red[t >> 6] = lo | (hi << 8);
  __syncthreads();
  lo = 255u;
  hi = 0u;
  for (int k = 0; k < 4; k++) {
    lo = (s_red[k] & 255u) < lo ? (s_red[k] & 255u) : lo;
    hi = (s_red[k] >> 8) > hi ? (s_red[k] >> 8) : hi;
  }
  __syncthreads();
  *mn = (float)lo;
  *mx = (float)hi;
}

// deblock.c:69-130, one workgroup per chunk. nbr[c * 6 + f] indexes chunks
// (faces xlo, xhi, ylo, yhi, zlo, zhi); out of range means no neighbour.
__global__ __launch_bounds__(256) void k_db_chunks(const uint8_t *chunks, const int32_t *nbr, uint32_t n, float step,
                                                   float strength, uint8_t *out) {
  __shared__ __align__(16) uint8_t s[4096];
  __shared__ uint32_t s_red[4];
  const uint32_t t = threadIdx.x;
  const uint32_t a = t >> 4, b = t & 15u;
  for (uint32_t c = blockIdx.x; c < n; c += gridDim.x) {
    const uint8_t *ctr = chunks + (size_t)c * 4096u;
    ((uint4 *)s)[t] = ((const uint4 *)ctr)[t];
    __syncthreads();
    if (strength > 0.0f) {
      float cmn = 0.0f, cmx = 0.0f;
      if (!(step > 0.0f)) block_minmax(ctr, t, s_red, &cmn, &cmx);
      for (uint32_t f = 0; f < 6; f++) {
        const int32_t j = nbr[(size_t)c * 6u + f];
        if (j < 0 || (uint32_t)j >= n) continue;  // the whole workgroup agrees
        const uint8_t *nb = chunks + (size_t)j * 4096u;
        float range = 1.0f;
        if (!(step > 0.0f)) {
          float nmn, nmx;
          block_minmax(nb, t, s_red, &nmn, &nmx);
          range = (cmx > nmx ? cmx : nmx) - (cmn < nmn ? cmn : nmn);
        }
        if (range > 0.0f) {
          float alpha, beta;
          db_limits(step, range, strength, &alpha, &beta);
          const bool lo = (f & 1u) == 0u;
          const uint32_t axis = f >> 1;
          const uint32_t c0d = lo ? 0u : 15u, c1d = lo ? 1u : 14u, n0d = lo ? 15u : 0u, n1d = lo ? 14u : 1u;
          auto at = [&](uint32_t d) {  // voxel at depth d from the face, in-plane (a, b)
            return axis == 0 ? (a * 16u + b) * 16u + d : axis == 1 ? (a * 16u + d) * 16u + b : (d * 16u + a) * 16u + b;
          };
          const uint32_t ci0 = at(c0d), ci1 = at(c1d);
          const float n0 = nb[at(n0d)], n1 = nb[at(n1d)], c0 = s[ci0], c1 = s[ci1];
          const float p1 = lo ? n1 : c1, p0 = lo ? n0 : c0, q0 = lo ? c0 : n0, q1 = lo ? c1 : n1;
          float d1;
          const float d0 = db_delta(p1, p0, q0, q1, alpha, beta, &d1);
          if (d0 != 0.0f) {
            s[ci0] = db_u8(lo ? c0 - d0 : c0 + d0);
            s[ci1] = db_u8(lo ? c1 - d1 : c1 + d1);
          }
        }
        __syncthreads();
      }
    }
    ((uint4 *)(out + (size_t)c * 4096u))[t] = ((const uint4 *)s)[t];
    __syncthreads();
  }
}

}  // namespace r3d

using namespace r3d;

extern "C" int r3d_hip_deblock_u8(r3d_hip_ctx *ctx, uint8_t *vol, const uint64_t dims[3], float step, float strength,
                                  void *stream) {
  if (!ctx) return R3D_ERR_INVALID;
  if (!dims) return r3d_fail(ctx, R3D_ERR_INVALID, "deblock: NULL dims");
  const uint64_t nz = dims[0], ny = dims[1], nx = dims[2];
  if (nz == 0 || ny == 0 || nx == 0 || !(strength > 0.0f)) return R3D_OK;  // deblock.c:137-139: no-op
  if (!vol) return r3d_fail(ctx, R3D_ERR_INVALID, "deblock: NULL volume");
  if (ny > (1ull << 40) / nx || nz > (1ull << 40) / (ny * nx))
    return r3d_fail(ctx, R3D_ERR_RANGE, "deblock: volume too large");
  R3D_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  hipStream_t st = (hipStream_t)stream;
  DbPass p{};
  p.vol = vol;
  p.nz = nz;
  p.ny = ny;
  p.nx = nx;
  if (step > 0.0f) {
    db_limits(step, 0.0f, strength, &p.lim.alpha, &p.lim.beta);
    p.lim.on = 1u;
    p.flim = nullptr;
  } else {
    // the context's own 16 bytes, not the shard workspace: a region decode in
    // flight keeps its parsed levels there
    static_assert(sizeof(DbLim) <= 16u, "DbLim fits the context's limits word");
    p.flim = (const DbLim *)ctx->d_dblim;
    hipLaunchKernelGGL(k_db_range, dim3(1), dim3(1024), 0, st, vol, nz * ny * nx, strength, (DbLim *)ctx->d_dblim);
  }
  const uint64_t seams[3] = {nx >= 2u ? (nx - 2u) / DB_N : 0u, ny >= 2u ? (ny - 2u) / DB_N : 0u,
                             nz >= 2u ? (nz - 2u) / DB_N : 0u};
  if (nx % 16u == 0u && (uintptr_t)vol % 16u == 0u && nx <= 0xFFFFFFF0u && ny <= 0xFFFFFFF0u && nz <= 0xFFFFFFF0u) {
    DbTile tl{};
    tl.vol = vol;
    tl.nx = (uint32_t)nx;
    tl.ny = (uint32_t)ny;
    tl.nz = (uint32_t)nz;
    tl.tx = (uint32_t)((nx + 2u + DBT_W - 1u) / DBT_W);
    tl.ty = (uint32_t)((ny + 2u + 15u) / 16u);
    tl.tz = (uint32_t)((nz + 2u + 15u) / 16u);
    tl.lim = p.lim;
    tl.flim = p.flim;
    const uint64_t tiles = (uint64_t)tl.tx * tl.ty * tl.tz;
    if (tiles <= 0x7FFFFFFFull) {
      hipLaunchKernelGGL(k_db_tiles<DBT_W>, dim3((uint32_t)tiles), dim3(256), 0, st, tl);
      R3D_HIP_CHECK(ctx, hipGetLastError());
      return R3D_OK;
    }
  }
  // vector width: rows aligned to 16 (or 4) bytes take 16 (or 4) samples per lane
  const uintptr_t base = (uintptr_t)vol;
  const int w = (nx % 16u == 0u && base % 16u == 0u) ? 16 : (nx % 4u == 0u && base % 4u == 0u) ? 4 : 1;
  const uint64_t totals[3] = {nz * ny * seams[0], nz * seams[1] * (nx / (uint64_t)w),
                              seams[2] * ny * (nx / (uint64_t)w)};
  void (*kern[3])(DbPass);
  kern[0] = w >= 4 ? k_db_pass<0, 4> : k_db_pass<0, 1>;
  kern[1] = w == 16 ? k_db_pass<1, 16> : w == 4 ? k_db_pass<1, 4> : k_db_pass<1, 1>;
  kern[2] = w == 16 ? k_db_pass<2, 16> : w == 4 ? k_db_pass<2, 4> : k_db_pass<2, 1>;
  for (int ax = 0; ax < 3; ax++) {
    if (!totals[ax]) continue;
    p.nseam = seams[ax];
    p.total = totals[ax];
    const uint64_t blocks = (p.total + 255u) / 256u;
    const uint32_t grid = (uint32_t)(blocks < (uint64_t)ctx->num_cus * 32u ? blocks : (uint64_t)ctx->num_cus * 32u);
    hipLaunchKernelGGL(kern[ax], dim3(grid ? grid : 1u), dim3(256), 0, st, p);
  }
  R3D_HIP_CHECK(ctx, hipGetLastError());
  return R3D_OK;
}

extern "C" int r3d_hip_deblock_chunks_u8(r3d_hip_ctx *ctx, uint32_t nchunks, const uint8_t *chunks,
                                         const int32_t *neighbors, float step, float strength, uint8_t *out,
                                         void *stream) {
  if (!ctx) return R3D_ERR_INVALID;
  if (nchunks == 0) return R3D_OK;
  if (!chunks || !neighbors || !out) return r3d_fail(ctx, R3D_ERR_INVALID, "deblock_chunks: NULL argument");
  if (((uintptr_t)chunks & 15u) || ((uintptr_t)out & 15u))
    return r3d_fail(ctx, R3D_ERR_INVALID, "deblock_chunks: chunks and out must be 16-byte aligned");
  if ((const uint8_t *)out < chunks + (size_t)nchunks * 4096u && chunks < (const uint8_t *)out + (size_t)nchunks * 4096u)
    return r3d_fail(ctx, R3D_ERR_INVALID, "deblock_chunks: out overlaps chunks");
  R3D_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  const uint32_t cap = (uint32_t)(ctx->num_cus > 0 ? ctx->num_cus : 256) * 16u;
  hipLaunchKernelGGL(k_db_chunks, dim3(nchunks < cap ? nchunks : cap), dim3(256), 0, (hipStream_t)stream, chunks,
                     neighbors, nchunks, step, strength, out);
  R3D_HIP_CHECK(ctx, hipGetLastError());
  return R3D_OK;
}
