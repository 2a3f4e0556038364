// r3d_decode.hip -- 3ddct chunk decode on gfx950.
//
// k_decode_exact follows the pinned FP32 sequence of the reference decoder
// (tools/3ddct/dct3d.c) with no fused multiply-add (the library is compiled
// with -ffp-contract=off):
//   coef = deq_one(level, step[cz+cy+cx])                  dct3d.c:149-154
//   three rounds of (16-point DCT-III on 256 lines, rotate (z,y,x)->(x,z,y))
//                                                          dct3d.c:61-118
//     per line: e[n] = sum over even k ascending of cm[k][n]*in[k],
//               o[n] likewise over odd k, zero inputs skipped;
//               out[n] = e[n] + o[n], out[15-n] = e[n] - o[n]
//   norm = rb + dcq/16, corrections in order, v = norm*(vspan/255) + vmin
//                                                          dct3d.c:568-591
//   u8 = clamp(roundf(v), 0, 255), NaN -> 0                dct3d.c:625-635
//
// Layout: one 256-lane workgroup per chunk at a time (at most 56 workgroups
// per CU walk the chunks with a grid stride), lane t owns line t
// of the current pass in registers; the rotate between passes is a
// transposing write into a 16 KiB LDS tile (lane t writes element c of its
// line to tile[c*256 + t], so the next pass's line t' = tile[t'*16 .. +15]).
// Skipping a zero input only drops an exact +-0 term, so a lane whose wave
// neighbours still run that k-term gets the same sums as the reference.
#include <math.h>

#include "r3d_dmath.h"
#include "r3d_internal.h"

namespace r3d {

__constant__ float c_basis[256];
__constant__ float c_mag[16];
constexpr int basis_j(int k, int n);
constexpr bool basis_neg(int k, int n);

int upload_basis(r3d_hip_ctx *ctx) {
  // magnitudes by fold index, then a bit-exact check of every entry
  float mag[16] = {0};
  for (int k = 0; k < 16; k++)
    for (int n = 0; n < 8; n++) mag[basis_j(k, n)] = __builtin_fabsf(ctx->basis[k * 16 + n]);
  for (int k = 0; k < 16; k++)
    for (int n = 0; n < 8; n++) {
      const float want = basis_neg(k, n) ? -mag[basis_j(k, n)] : mag[basis_j(k, n)];
      if (__builtin_memcmp(&want, &ctx->basis[k * 16 + n], 4) != 0)
        return r3d_fail(ctx, R3D_ERR_DEVICE, "DCT basis entry (%d,%d) breaks the cosine symmetry", k, n);
    }
  R3D_HIP_CHECK(ctx, hipMemcpyToSymbol(HIP_SYMBOL(c_basis), ctx->basis, sizeof ctx->basis, 0,
                                       hipMemcpyHostToDevice));
  R3D_HIP_CHECK(ctx, hipMemcpyToSymbol(HIP_SYMBOL(c_mag), mag, sizeof mag, 0, hipMemcpyHostToDevice));
  return R3D_OK;
}

struct DecodeArgs {
  uint32_t nchunks, fmt;
  const int16_t *levels;
  const uint32_t *pair_off;
  const uint32_t *pairs;
  uint64_t npairs;
  const r3d_chunk_header *hdr;
  const float *tables;
  uint32_t ntables;
  const uint32_t *corr_off;
  const uint16_t *corr_pos;
  const float *corr_delta;
  uint64_t ncorr;
  uint8_t *out;
  const uint64_t *out_off;
  uint32_t sy, sz;
  int32_t *status;
  const uint32_t *pair_cnt;
  const uint32_t *corr_cnt;
  // Region destination (shard region decodes, exact kernel only): chunk c is
  // cell rfirst + c of a chunk box (x fastest, ncy x ncx per slice) whose first chunk
  // starts at (rz, ry, rx) relative to the region origin. Voxels outside
  // [0, hz) x [0, hy) x [0, hx) of the region are not written; the region is
  // x fastest with row stride rsy and slice stride rsz.
  uint32_t rgn, ncy, ncx, rfirst;  // rfirst: the box cell of chunk 0
  int64_t rz, ry, rx, hz, hy, hx;
  uint64_t rsy, rsz;
};

// [begin, end) of chunk c's pairs: offsets are caller data, so the range is
// clamped to the next offset, to the count when given and to npairs.
__device__ __forceinline__ void pair_range(const DecodeArgs &a, uint32_t c, uint64_t *b, uint64_t *e) {
  uint64_t p0 = a.pair_off[c], p1 = a.pair_off[c + 1];
  if (a.pair_cnt && p0 + a.pair_cnt[c] < p1) p1 = p0 + a.pair_cnt[c];
  if (p1 > a.npairs) p1 = a.npairs;
  *b = p0;
  *e = p1;
}

// Chunk c's corrections, clamped the same way and to 4096.
__device__ __forceinline__ void corr_range(const DecodeArgs &a, uint32_t c, uint64_t *b, uint64_t *n) {
  const uint64_t c0 = a.corr_off[c];
  uint64_t c1 = a.corr_off[c + 1] < a.ncorr ? a.corr_off[c + 1] : a.ncorr;
  if (a.corr_cnt && c0 + a.corr_cnt[c] < c1) c1 = c0 + a.corr_cnt[c];
  *b = c0;
  *n = c1 > c0 ? (c1 - c0 > 4096u ? 4096u : c1 - c0) : 0;
}

// deq_one (dct3d.c:149-154), r = (|l| - 1) * step + 0.80f * step + 0.40f * step
// with 0.80f * step and 0.40f * step taken from a table (the same rounded
// products, computed once per table entry). The sign and the zero case are
// one product: r > 0 for every step > 0, and r * med3(l, -1, 1) is exactly
// -r, +0 or r, as the reference's branches give.
__device__ __forceinline__ float deq_tab(float lv, float step, float s8, float s4) {
  const float r = (__builtin_fabsf(lv) - 1.0f) * step + s8 + s4;
  return r * fclamp_ord(lv, -1.0f, 1.0f);
}

// dct3d.c:625-635 for the u8 output: clamp(roundf(v), 0, 255), NaN -> 0.
// With c = 0.49999997f (the float just below 0.5), trunc(v + c), the sum in
// FP32 rounded to nearest even, is roundf(v) for every float v >= 0: a tie
// k + 0.5 sums to within half an ulp of k + 1 and rounds up to it, and
// anything below a tie stays below k + 1 (c < 0.5 keeps 0.5 - 2^-25 at 0).
// Clamping the sum to [0, 255] then gives 0 for v < 0 and NaN (maxnum) and
// 255 above. Identical to the reference's for every float (checked
// exhaustively over [-1, 65536]; beyond, both saturate), in 3-4 instructions
// instead of 5.
__device__ __forceinline__ uint32_t round_u8(float v) {
  const float s = __builtin_fmaxf(v + 0.49999997f, 0.0f);
  return (uint32_t)__builtin_fminf(s, 255.0f);
}

// The basis has 15 distinct magnitudes: cm[k][n] = +-M[j] with
// j = |((2n+1)k mod 64)| folded to [0,16) (j = 0 is c_0 = 1/4, equal to
// M[8]) and the sign negative for (2n+1)k mod 64 in (16, 48). The host
// checks that the FP32 table it built reproduces this structure bit for bit
// (upload_basis), so (-M)*v == cm*v exactly and e - M*v == e + cm*v.
constexpr int basis_j(int k, int n) {
  return k == 0 ? 0 : (((2 * n + 1) * k) & 31) < 32 - (((2 * n + 1) * k) & 31)
                          ? (((2 * n + 1) * k) & 31)
                          : 32 - (((2 * n + 1) * k) & 31);
}
constexpr bool basis_neg(int k, int n) {
  return (((2 * n + 1) * k) & 63) > 16 && (((2 * n + 1) * k) & 63) < 48;
}

// dct3d.c:52-59 (forward, for the workload encoder).
__device__ __forceinline__ void fdct16(const float in[16], float out[16]) {
  float s[8], d[8];
#pragma unroll
  for (int n = 0; n < 8; n++) {
    s[n] = in[n] + in[15 - n];
    d[n] = in[n] - in[15 - n];
  }
#pragma unroll
  for (int k = 0; k < 16; k += 2) {
    float a = 0.0f;
#pragma unroll
    for (int n = 0; n < 8; n++) a += c_basis[k * 16 + n] * s[n];
    out[k] = a;
  }
#pragma unroll
  for (int k = 1; k < 16; k += 2) {
    float a = 0.0f;
#pragma unroll
    for (int n = 0; n < 8; n++) a += c_basis[k * 16 + n] * d[n];
    out[k] = a;
  }
}

// Where row (z, y) of chunk c goes, and which of its 16 voxels are stored.
struct RowDst {
  uint8_t *p;
  uint32_t lx, hx;
  bool on;
};
__device__ __forceinline__ RowDst row_dst(const DecodeArgs &a, uint32_t c, uint32_t z, uint32_t y) {
  RowDst d;
  if (!a.rgn) {
    d.p = a.out + (a.out_off ? a.out_off[c] : (uint64_t)c * 4096u) + z * a.sz + y * a.sy;
    d.lx = 0;
    d.hx = 16;
    d.on = true;
    return d;
  }
  const uint32_t cl = a.rfirst + c;
  const uint32_t ix = cl % a.ncx, r = cl / a.ncx, iy = r % a.ncy, iz = r / a.ncy;
  const int64_t gz = a.rz + 16 * (int64_t)iz + z, gy = a.ry + 16 * (int64_t)iy + y;
  const int64_t ox = a.rx + 16 * (int64_t)ix;
  d.on = gz >= 0 && gz < a.hz && gy >= 0 && gy < a.hy;
  const int64_t lx = ox < 0 ? -ox : 0, hx = a.hx - ox < 16 ? a.hx - ox : 16;
  d.lx = (uint32_t)lx;
  d.hx = hx > lx ? (uint32_t)hx : (uint32_t)lx;
  // byte address of the row's x = 0, which may lie before the region when
  // ox < 0; only x in [lx, hx) is dereferenced
  d.p = (uint8_t *)((uintptr_t)a.out + (uint64_t)(gz * (int64_t)a.rsz + gy * (int64_t)a.rsy + ox));
  return d;
}

__device__ __forceinline__ void store_row16(uint8_t *dst, const uint32_t w[4]);
__device__ __forceinline__ void store_row(const RowDst &d, const uint32_t w[4]) {
  if (!d.on) return;
  if (d.lx == 0 && d.hx == 16) {
    store_row16(d.p, w);
    return;
  }
  for (uint32_t x = d.lx; x < d.hx; x++) d.p[x] = (uint8_t)(w[x >> 2] >> ((x & 3) * 8));
}

__device__ __forceinline__ void store_row16(uint8_t *dst, const uint32_t w[4]) {
  const uintptr_t a = (uintptr_t)dst;
  if ((a & 15u) == 0) {
    *(uint4 *)dst = make_uint4(w[0], w[1], w[2], w[3]);
  } else if ((a & 3u) == 0) {
#pragma unroll
    for (int i = 0; i < 4; i++) ((uint32_t *)dst)[i] = w[i];
  } else {
#pragma unroll
    for (int i = 0; i < 16; i++) dst[i] = (uint8_t)(w[i >> 2] >> ((i & 3) * 8));
  }
}

// dct3d.c:61-68 on a register-resident line, over the inputs k <= K only (K
// uniform across the workgroup; a compile-time 15 for dense chunks): the
// inputs past K are zero, so their terms are exact +-0 that the reference
// skips, and here they may hold anything. Every term up to K is added (a
// zero input contributes an exact +-0 too), each sum in ascending k, and each
// sum starts from its first term instead of 0 + term: both differ from the
// reference only in the sign of an exact zero, which the + dc after the last
// pass removes.
__device__ __forceinline__ void idct16k(const float in[16], float out[16], uint32_t K) {
  float e[8], o[8];
#pragma unroll
  for (int n = 0; n < 8; n++) {
    const float p = c_mag[basis_j(0, n)] * in[0];
    e[n] = basis_neg(0, n) ? -p : p;
  }
  if (K >= 1u) {
#pragma unroll
    for (int n = 0; n < 8; n++) {
      const float q = c_mag[basis_j(1, n)] * in[1];
      o[n] = basis_neg(1, n) ? -q : q;
    }
  } else {
#pragma unroll
    for (int n = 0; n < 8; n++) o[n] = 0.0f;
  }
#pragma unroll
  for (int k = 2; k < 16; k++) {
    if ((uint32_t)k <= K) {
      const float v = in[k];
#pragma unroll
      for (int n = 0; n < 8; n++) {
        const float p = c_mag[basis_j(k, n)] * v;
        if (k & 1)
          o[n] = basis_neg(k, n) ? o[n] - p : o[n] + p;
        else
          e[n] = basis_neg(k, n) ? e[n] - p : e[n] + p;
      }
    }
  }
#pragma unroll
  for (int n = 0; n < 8; n++) {
    out[n] = e[n] + o[n];
    out[15 - n] = e[n] - o[n];
  }
}

// The wave-wide maximum of v in [0, 31] (a binary search over ballots).
__device__ __forceinline__ uint32_t wave_max5(uint32_t v) {
  uint32_t m = 0;
#pragma unroll
  for (uint32_t b = 16; b; b >>= 1)
    if (__ballot(v >= m + b)) m += b;
  return m;
}

// A chunk's frequency box, packed per wave as (kx+1) | (ky+1) << 5 | (kz+1) << 10
// (0 where the wave saw no nonzero level).
__device__ __forceinline__ uint32_t box_pack(uint32_t xm, uint32_t ym, uint32_t zm) {
  return xm | ym << 5 | zm << 10;
}

// SPARSE: the pairs format, with the passes bounded by the chunk's frequency
// box; dense levels take every pass whole (their box is nearly always full).
// Held to 7 waves per SIMD (the one-product dequantisation otherwise takes
// the dense instance to 92 VGPRs, 5 waves).
template <bool F32OUT, bool SPARSE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(7))) void k_decode_exact(DecodeArgs a) {
  __shared__ float tile[4096];
  __shared__ float steps[3][48];  // step, 0.80f * step, 0.40f * step
  __shared__ uint32_t s_box[4];
  const uint32_t t = threadIdx.x, wv = t >> 6, lane = t & 63u;
  const uint32_t z = t >> 4, y = t & 15u;
  for (uint32_t c = blockIdx.x; c < a.nchunks; c += gridDim.x) {
    const r3d_chunk_header h = a.hdr[c];
    const float dc = (float)h.dcq / 16.0f;
    const bool ok = h.table < a.ntables && h.vspan > 0.0f && __builtin_isfinite(h.vmin) &&
                    __builtin_isfinite(dc);
    const RowDst dst = row_dst(a, c, z, y);
    if (!ok) {  // dct3d.c:555-556: rejected header; the chunk is zero-filled
      const uint32_t zw[4] = {0, 0, 0, 0};
      if (F32OUT) {
        float4 *fd = (float4 *)(a.out + ((size_t)c * 4096u + t * 16u) * 4u);
        for (int i = 0; i < 4; i++) fd[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      } else {
        store_row(dst, zw);
      }
      if (a.status && t == 0) a.status[c] = 0;
      continue;
    }
    if (t < 46) {
      const float st = a.tables[(size_t)h.table * 46u + t];
      steps[0][t] = st;
      steps[1][t] = 0.80f * st;
      steps[2][t] = 0.40f * st;
    }
    // the frequency box: every nonzero level has x <= kx, y <= ky, z <= kz
    int32_t w[8];
    uint64_t p0 = 0, p1 = 0;
    uint32_t kx = 15u, ky = 15u, kz = 15u;
    if (!SPARSE) {
      const int4 *src = (const int4 *)(a.levels + (size_t)c * 4096u + t * 16u);
      const int4 q0 = src[0], q1 = src[1];
      w[0] = q0.x, w[1] = q0.y, w[2] = q0.z, w[3] = q0.w, w[4] = q1.x, w[5] = q1.y, w[6] = q1.z, w[7] = q1.w;
      __syncthreads();
    } else {
      pair_range(a, c, &p0, &p1);
      uint32_t xm = 0, ym = 0, zm = 0;
      for (uint64_t p = p0 + t; p < p1; p += 256u) {
        const uint32_t pos = a.pairs[p] & 4095u;
        xm = max(xm, (pos & 15u) + 1u);
        ym = max(ym, ((pos >> 4) & 15u) + 1u);
        zm = max(zm, (pos >> 8) + 1u);
      }
      const uint32_t bx = wave_max5(xm), by = wave_max5(ym), bz = wave_max5(zm);
      if (lane == 0) s_box[wv] = box_pack(bx, by, bz);
      __syncthreads();
      uint32_t bxm = 0, bym = 0, bzm = 0;
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const uint32_t b = __builtin_amdgcn_readfirstlane(s_box[i]);
        bxm = max(bxm, b & 31u);
        bym = max(bym, (b >> 5) & 31u);
        bzm = max(bzm, b >> 10);
      }
      kx = bxm ? bxm - 1u : 0u, ky = bym ? bym - 1u : 0u, kz = bzm ? bzm - 1u : 0u;
    }
    const bool p1_wave = 4u * wv <= kz;  // pass 1: the waves holding lines z <= kz
    float v[16], o[16];
    if (!SPARSE) {
#pragma unroll
      for (int k = 0; k < 16; k++) {
        const int16_t l = (int16_t)((uint32_t)w[k >> 1] >> ((k & 1) * 16));
        const uint32_t f = z + y + k;
        v[k] = deq_tab((float)l, steps[0][f], steps[1][f], steps[2][f]);
      }
    } else {
      // zeros for the box's rows, then the pairs
      if (z <= kz && y <= ky) {
        float4 *row = (float4 *)&tile[t * 16u];
#pragma unroll
        for (int i = 0; i < 4; i++) row[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      }
      __syncthreads();
      for (uint64_t p = p0 + t; p < p1; p += 256u) {
        const uint32_t pr = a.pairs[p];
        const uint32_t pos = pr & 4095u;
        const float l = (float)(int16_t)(pr >> 16);
        const uint32_t f = (pos >> 8) + ((pos >> 4) & 15u) + (pos & 15u);
        tile[pos] = deq_tab(l, steps[0][f], steps[1][f], steps[2][f]);
      }
      __syncthreads();
      if (p1_wave) {
#pragma unroll
        for (int k = 0; k < 16; k++) v[k] = tile[t * 16 + k];
      }
      __syncthreads();
    }
    // pass 1 (along x): line t = (z, y); writes tile[x*256 + t] (lines outside
    // the box hold anything: the next pass reads none of them)
    if (p1_wave) {
      idct16k(v, o, kx);
#pragma unroll
      for (int k = 0; k < 16; k++) tile[k * 256 + t] = o[k];
    }
    __syncthreads();
    // pass 2 (along y): the lines (x, z <= kz) packed into the first lanes
    const uint32_t d2 = kz + 1u, m2 = (65536u + d2 - 1u) / d2;
    const uint32_t x2 = (t * m2) >> 16, z2 = t - x2 * d2;
    const bool p2 = t < 16u * d2;
    if (p2) {
#pragma unroll
      for (int k = 0; k < 16; k++) v[k] = tile[x2 * 256u + z2 * 16u + k];
    }
    __syncthreads();
    if (p2) {
      idct16k(v, o, ky);
#pragma unroll
      for (int k = 0; k < 16; k++) tile[k * 256 + x2 * 16u + z2] = o[k];
    }
    __syncthreads();
    // pass 3 (along z): every line t = (y, x); writes natural order + dc
#pragma unroll
    for (int k = 0; k < 16; k++) v[k] = tile[t * 16 + k];
    __syncthreads();
    idct16k(v, o, kz);
#pragma unroll
    for (int k = 0; k < 16; k++) tile[k * 256 + t] = o[k] + dc;
    __syncthreads();
    if (a.corr_off) {  // dct3d.c:572-587: serial, in stream order
      if (t == 0) {
        uint64_t c0, n;
        corr_range(a, c, &c0, &n);
        for (uint64_t i = 0; i < n; i++) {
          const uint32_t pos = a.corr_pos[c0 + i];
          if (pos >= 4096u) break;
          tile[pos] += a.corr_delta[c0 + i];
        }
      }
      __syncthreads();
    }
    const float inv = h.vspan / 255.0f;
    if (F32OUT) {
      float4 *fd = (float4 *)(a.out + ((size_t)c * 4096u + t * 16u) * 4u);
#pragma unroll
      for (int x = 0; x < 16; x += 4)
        fd[x >> 2] = make_float4(tile[t * 16 + x] * inv + h.vmin, tile[t * 16 + x + 1] * inv + h.vmin,
                                 tile[t * 16 + x + 2] * inv + h.vmin, tile[t * 16 + x + 3] * inv + h.vmin);
    } else {
      uint32_t wo[4] = {0, 0, 0, 0};
#pragma unroll
      for (int x = 0; x < 16; x++) wo[x >> 2] = round_u8_into(tile[t * 16 + x] * inv + h.vmin, x & 3, wo[x >> 2]);
      store_row(dst, wo);
    }
    if (a.status && t == 0) a.status[c] = 1;
    __syncthreads();
  }
}

// ---- workload encoder (ENC_BODY + encode_float up to the quantiser) -----

struct EncodeArgs {
  const uint8_t *vol;
  uint32_t nx, ny, nz, cx, cy;
  uint32_t nchunks;
  const float *step;
  int16_t *levels;
  r3d_chunk_header *hdr;
};

__device__ __forceinline__ float quant_one(float c, float step) {
  float dz = 0.80f * step, av = __builtin_fabsf(c);
  float lv = 0.0f;
  if (av >= dz) lv = __builtin_floorf((av - dz) / step + 1.0f);
  return c < 0.0f ? -lv : lv;
}

__global__ __launch_bounds__(256) void k_encode(EncodeArgs a) {
  __shared__ float tile[4096];
  __shared__ float red[3][256];
  __shared__ float steps[48];
  const uint32_t t = threadIdx.x;
  const uint32_t z = t >> 4, y = t & 15u;
  if (t < 46) steps[t] = a.step[t];
  for (uint32_t c = blockIdx.x; c < a.nchunks; c += gridDim.x) {
    const uint32_t bx = c % a.cx, by = (c / a.cx) % a.cy, bz = c / (a.cx * a.cy);
    const uint8_t *row =
        a.vol + (((size_t)(bz * 16 + z) * a.ny + (by * 16 + y)) * a.nx + bx * 16);
    const uint4 q = *(const uint4 *)row;
    const uint32_t qw[4] = {q.x, q.y, q.z, q.w};
    float in[16];
    float mn = 255.0f, mx = 0.0f;
#pragma unroll
    for (int i = 0; i < 16; i++) {
      in[i] = (float)((qw[i >> 2] >> ((i & 3) * 8)) & 0xffu);
      mn = fmin_sel(mn, in[i]);
      mx = fmax_sel(mx, in[i]);
    }
    red[0][t] = mn;
    red[1][t] = mx;
    __syncthreads();
    for (uint32_t s = 128; s > 0; s >>= 1) {
      if (t < s) {
        red[0][t] = fmin_sel(red[0][t], red[0][t + s]);
        red[1][t] = fmax_sel(red[1][t], red[1][t + s]);
      }
      __syncthreads();
    }
    const float vmin = red[0][0], vmax = red[1][0];
    float vspan = vmax - vmin;
    if (!(vspan > 0.0f)) vspan = 1.0f;
    const float scale = 255.0f / vspan;
    float rs = 0.0f;
#pragma unroll
    for (int i = 0; i < 16; i++) {
      in[i] = (in[i] - vmin) * scale;
      rs += in[i];
    }
    __syncthreads();
    red[2][t] = rs;
    __syncthreads();
    for (uint32_t s = 128; s > 0; s >>= 1) {
      if (t < s) red[2][t] += red[2][t + s];
      __syncthreads();
    }
    const float dcf = red[2][0] / 4096.0f;
    const int32_t dcq_i = (int32_t)__builtin_roundf(dcf * 16.0f);
    const float dcq = (float)dcq_i / 16.0f;
#pragma unroll
    for (int i = 0; i < 16; i++) in[i] = in[i] - dcq;
    float o[16];
    fdct16(in, o);
#pragma unroll
    for (int k = 0; k < 16; k++) tile[k * 256 + t] = o[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 16; k++) in[k] = tile[t * 16 + k];
    __syncthreads();
    fdct16(in, o);
#pragma unroll
    for (int k = 0; k < 16; k++) tile[k * 256 + t] = o[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 16; k++) in[k] = tile[t * 16 + k];
    __syncthreads();
    fdct16(in, o);
#pragma unroll
    for (int k = 0; k < 16; k++) tile[k * 256 + t] = o[k];
    __syncthreads();
    int16_t *dst = a.levels + (size_t)c * 4096u + t * 16u;
    uint32_t pw[8];
#pragma unroll
    for (int k = 0; k < 16; k += 2) {
      float l0 = quant_one(tile[t * 16 + k], steps[z + y + k]);
      float l1 = quant_one(tile[t * 16 + k + 1], steps[z + y + k + 1]);
      l0 = fclamp_sel(l0, -32768.0f, 32767.0f);
      l1 = fclamp_sel(l1, -32768.0f, 32767.0f);
      pw[k >> 1] = (uint32_t)(uint16_t)(int16_t)l0 | ((uint32_t)(uint16_t)(int16_t)l1 << 16);
    }
    ((uint4 *)dst)[0] = make_uint4(pw[0], pw[1], pw[2], pw[3]);
    ((uint4 *)dst)[1] = make_uint4(pw[4], pw[5], pw[6], pw[7]);
    if (t == 0) {
      r3d_chunk_header hh;
      hh.vmin = vmin;
      hh.vspan = vspan;
      hh.dcq = dcq_i;
      hh.table = 0;
      a.hdr[c] = hh;
    }
    __syncthreads();
  }
}


// ---- MFMA variants (approximate) --------------------------------------------
//
// One wave per chunk, four independent waves per 256-lane workgroup,
// grid-strided. Lane l = r + 16 g (r = l % 16, g = l / 16). The 16 x 16 MFMA
// operand layouts used: A: lane l holds row r of A at its k set; B: lane l holds
// column r of B at its k set; D: lane l holds D[4g + e][r], e = 0..3. For
// v_mfma_f32_16x16x16_f16 the k set of lane l is 4g .. 4g+3 (slot e);
// v_mfma_f32_16x16x4_f32 takes one k per lane (k = g), and four instructions
// e = 0..3 cover contraction index 4g + e: the order in which the k terms are
// summed is free, so "slot e of every lane" is one instruction's k.
//
//   load    lane (r, g) holds X[z = 4g+e][y = r][x] for x = 0..15
//   pass 1  per x: D1 = X_x (A: row y, k = z) * C (B: k = z, column z')
//           -> lane (r, g) slot e = D1[y = 4g+e][z' = r]
//   pass 2  per x: D2 = C^T (A: row y', k = y) * D1 (B, exactly as pass 1 left it)
//           -> lane (r, g) slot e = D2[y' = 4g+e][z' = r]
//   pass 3  in registers: lane (r, g) holds all 16 x-frequencies of the four
//           rows (z' = r, y' = 4g+e) and runs an FP32 16-point inverse DCT
//           on each, so it ends with four whole 16-byte output rows.
//
// No LDS is needed for dense levels. Precision per variant:
//   PREC_F32   FP32 MFMA (v_mfma_f32_16x16x4_f32), FP32 throughout
//   PREC_F16   FP16 operands, FP32 accumulate (about 11 significant bits)
//   PREC_F16X3 each product as hi*hi + lo*hi + hi*lo of FP16 hi/lo halves
//              (hi rounded toward zero, lo the FP32 remainder; about 21 bits)
// The FP16 variants scale the chunk by a power of two chosen from its largest
// level and the table's largest step, so that every FP16 operand stays below
// 2^15 < 65504 (each pass of an orthonormal basis grows magnitudes at most
// 4x); the scale is undone in the epilogue. PREC_F16X3 carries the low halves
// times 2^10 so that they stay out of FP16's subnormal range.
// Dequantisation uses (|l| + 0.2) * step, equal in exact arithmetic to
// deq_one's (|l| - 1) * step + 0.8 * step + 0.4 * step.

typedef _Float16 h4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef short s2 __attribute__((ext_vector_type(2)));

enum : int { PREC_F32 = 1, PREC_F16 = 2, PREC_F16X3 = 3 };

__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax_sel(v, __shfl_xor(v, o, 64));
  return v;
}

// Orders this wave's LDS accesses (lanes of one wave exchanging data).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// power of two p with bound * p < 2^top; 1 when bound == 0
__device__ __forceinline__ float pow2_scale(float bound, int top) {
  if (!(bound > 0.0f)) return 1.0f;
  const int e = (int)((__float_as_uint(bound) >> 23) & 0xffu) - 127;  // floor(log2 bound), bound normal
  int k = top - 1 - e;
  k = k < -120 ? -120 : (k > 120 ? 120 : k);
  return __uint_as_float((uint32_t)(k + 127) << 23);
}

__device__ __forceinline__ h4 to_h4(const float v[4]) {
  h4 h;
#pragma unroll
  for (int e = 0; e < 4; e++) h[e] = (_Float16)v[e];
  return h;
}

// hi/lo split for PREC_F16X3: hi = v rounded toward zero to FP16 by one
// intrinsic (so the value the matrix core sees is the value lo is taken
// from), lo = (v - hi) * LO_SCALE. |v - hi| < 2^-10 |hi|, so the scaled low
// half stays in FP16's normal range wherever hi does (the matrix cores flush
// FP16 subnormal operands, which would drop an unscaled low half) and below
// 2^15 for |hi| < 2^15.
constexpr float LO_SCALE = 1024.0f, LO_UNSCALE = 1.0f / 1024.0f;
__device__ __forceinline__ void split_h4(const float v[4], h4 *hi, h4 *lo) {
  const uint2 u = make_uint2(__builtin_bit_cast(uint32_t, __builtin_amdgcn_cvt_pkrtz(v[0], v[1])),
                             __builtin_bit_cast(uint32_t, __builtin_amdgcn_cvt_pkrtz(v[2], v[3])));
  *hi = __builtin_bit_cast(h4, u);
#pragma unroll
  for (int e = 0; e < 4; e++) (*lo)[e] = (_Float16)((v[e] - (float)(*hi)[e]) * LO_SCALE);
}

// D = A * B on the MFMA layouts above; a and b hold this lane's four slots.
template <int PREC>
__device__ __forceinline__ f4 mma(const float a[4], const float b[4], h4 ah, h4 al, h4 bh, h4 bl) {
  f4 d = {0.0f, 0.0f, 0.0f, 0.0f};
  if (PREC == PREC_F32) {
#pragma unroll
    for (int e = 0; e < 4; e++) d = __builtin_amdgcn_mfma_f32_16x16x4f32(a[e], b[e], d, 0, 0, 0);
  } else {
    d = __builtin_amdgcn_mfma_f32_16x16x16f16(ah, bh, d, 0, 0, 0);
    if (PREC == PREC_F16X3) {  // hi*hi + 2^-10 (lo'*hi + hi*lo')
      f4 x = {0.0f, 0.0f, 0.0f, 0.0f};
      x = __builtin_amdgcn_mfma_f32_16x16x16f16(al, bh, x, 0, 0, 0);
      x = __builtin_amdgcn_mfma_f32_16x16x16f16(ah, bl, x, 0, 0, 0);
#pragma unroll
      for (int e = 0; e < 4; e++) d[e] = __builtin_fmaf(x[e], LO_UNSCALE, d[e]);
    }
  }
  return d;
}

#define R3D_CM(k, n) (basis_neg(k, n) ? -c_mag[basis_j(k, n)] : c_mag[basis_j(k, n)])

// 16-point inverse DCT (out[n] = sum_k cm[k][n] in[k]) as a partial butterfly
// with fused multiply-adds: odd k (8 x 8), k = 2 mod 4 (4 x 4), k = 4, 12 and
// k = 0, 8 (2 x 2 each), then the even/odd recombinations; 116 operations.
__device__ __forceinline__ void idct16_fma(const float in[16], float out[16]) {
  float o[8], eo[4], ee[4], e[8];
#pragma unroll
  for (int n = 0; n < 8; n++) {
    float acc = R3D_CM(1, n) * in[1];
#pragma unroll
    for (int k = 3; k < 16; k += 2) acc = __builtin_fmaf(R3D_CM(k, n), in[k], acc);
    o[n] = acc;
  }
#pragma unroll
  for (int n = 0; n < 4; n++) {
    float acc = R3D_CM(2, n) * in[2];
#pragma unroll
    for (int k = 6; k < 16; k += 4) acc = __builtin_fmaf(R3D_CM(k, n), in[k], acc);
    eo[n] = acc;
  }
#pragma unroll
  for (int n = 0; n < 2; n++) {
    const float eee = __builtin_fmaf(R3D_CM(8, n), in[8], R3D_CM(0, n) * in[0]);
    const float eeo = __builtin_fmaf(R3D_CM(12, n), in[12], R3D_CM(4, n) * in[4]);
    ee[n] = eee + eeo;
    ee[3 - n] = eee - eeo;
  }
#pragma unroll
  for (int n = 0; n < 4; n++) {
    e[n] = ee[n] + eo[n];
    e[7 - n] = ee[n] - eo[n];
  }
#pragma unroll
  for (int n = 0; n < 8; n++) {
    out[n] = e[n] + o[n];
    out[15 - n] = e[n] - o[n];
  }
}

// o[idx] += v for a wave-uniform idx, keeping o in registers.
__device__ __forceinline__ void add_at(float o[16], uint32_t idx, float v) {
  switch (idx) {
#define R3D_CASE(i) \
  case i:           \
    o[i] += v;      \
    break;
#define R3D_CASE8(i) R3D_CASE(i) R3D_CASE(i + 1) R3D_CASE(i + 2) R3D_CASE(i + 3) R3D_CASE(i + 4) \
    R3D_CASE(i + 5) R3D_CASE(i + 6) R3D_CASE(i + 7)
    R3D_CASE8(0) R3D_CASE8(8)
#undef R3D_CASE8
#undef R3D_CASE
    default:
      break;
  }
}

__device__ __forceinline__ float lane_level(const uint4 q[4][2], int e, int x) {
  const uint4 v = q[e][x >> 3];
  const uint32_t w = ((x & 7) >> 1) == 0 ? v.x : ((x & 7) >> 1) == 1 ? v.y : ((x & 7) >> 1) == 2 ? v.z : v.w;
  return (float)(int16_t)(w >> ((x & 1) * 16));
}

// Pass-3 LDS tile for PREC_F16, per wave: FP16 D2 rows (y', z') of 16
// x-frequencies at dword y' * P3_SY + z' * P3_SZ. With P3_SZ = 10 and
// P3_SY = 162 (both 2 mod 4 times an odd factor), the 64-bit writes (16 lanes
// of one y', z' = 0..15) and reads (16 lanes of one z', y' = 0..15) each
// touch 16 distinct even bank pairs.
constexpr uint32_t P3_SZ = 10, P3_SY = 162, P3_WAVE = 16 * P3_SY;

template <int PREC, bool SPARSE, bool F32OUT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(PREC == PREC_F16X3 ? 2 : 3))) void k_decode_mfma(
    DecodeArgs a) {
  constexpr bool MFMA3 = PREC == PREC_F16;  // pass 3 on MFMA through LDS
  // FP16 operand range: pass-1 inputs below 2^SCALE_TOP, pass-1 outputs at
  // most 4x that, pass-2 outputs 16x (converted to FP16 only for MFMA3)
  constexpr int SCALE_TOP = MFMA3 ? 11 : 13;
  // per wave: the sparse level tile (2048 words, used while loading) and the
  // pass-3 tile (used after the levels are in registers) share one region
  constexpr uint32_t WAVE_WORDS = (MFMA3 && P3_WAVE > 2048u) ? P3_WAVE : (SPARSE || MFMA3 ? 2048u : 4u);
  __shared__ __attribute__((aligned(16))) uint32_t lds_w[4 * WAVE_WORDS];
  const uint32_t lane = threadIdx.x & 63u, r = lane & 15u, g = lane >> 4;
  const uint32_t w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  float cb[4];  // cm[4g + e][r]: B of pass 1, A of passes 2 and 3
#pragma unroll
  for (int e = 0; e < 4; e++) cb[e] = c_basis[(4u * g + e) * 16u + r];
  h4 bh, bl = {};
  if (PREC == PREC_F16X3) split_h4(cb, &bh, &bl);
  else bh = to_h4(cb);
  // step[z + y + x] = step[(4g + r) + (e + x)]; reloaded when a chunk names
  // another table, with the table's largest step
  uint32_t cur_table = 0xffffffffu;
  float st_raw[19];
  float max_step = 0.0f;
  for (uint32_t c = blockIdx.x * 4u + w; c < a.nchunks; c += gridDim.x * 4u) {
    uint4 q[4][2];
    if (!SPARSE) {  // issued before the header is examined
      const int16_t *src = a.levels + (size_t)c * 4096u + r * 16u;
#pragma unroll
      for (int e = 0; e < 4; e++)
#pragma unroll
        for (int hh = 0; hh < 2; hh++) q[e][hh] = *(const uint4 *)(src + (4u * g + e) * 256u + hh * 8u);
    }
    const r3d_chunk_header h = a.hdr[c];
    const float dc = (float)h.dcq / 16.0f;
    const bool ok = h.table < a.ntables && h.vspan > 0.0f && __builtin_isfinite(h.vmin) && __builtin_isfinite(dc);
    uint8_t *dst = a.out + (a.out_off ? a.out_off[c] : (uint64_t)c * 4096u);
    float *fdst = (float *)a.out + (size_t)c * 4096u;
    if (!ok) {  // rejected header: zero-filled, status 0
#pragma unroll
      for (int e = 0; e < 4; e++) {
        if (F32OUT) {
#pragma unroll
          for (int i = 0; i < 4; i++)
            ((float4 *)(fdst + r * 256u + (4u * g + e) * 16u))[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        } else {
          const uint32_t zw[4] = {0, 0, 0, 0};
          store_row16(dst + r * a.sz + (4u * g + e) * a.sy, zw);
        }
      }
      if (a.status && lane == 0) a.status[c] = 0;
      continue;
    }
    if (SPARSE) {
      int16_t *t = (int16_t *)(lds_w + w * WAVE_WORDS);
#pragma unroll
      for (int i = 0; i < 8; i++) ((uint4 *)t)[lane + 64u * i] = make_uint4(0, 0, 0, 0);
      wave_sync();
      uint64_t p0, p1;
      pair_range(a, c, &p0, &p1);
      for (uint64_t p = p0 + lane; p < p1; p += 64u) {
        const uint32_t pr = a.pairs[p];
        t[pr & 4095u] = (int16_t)(pr >> 16);
      }
      wave_sync();
#pragma unroll
      for (int e = 0; e < 4; e++)
#pragma unroll
        for (int hh = 0; hh < 2; hh++) q[e][hh] = *(const uint4 *)(t + (4u * g + e) * 256u + r * 16u + hh * 8u);
      wave_sync();
    }
    if (h.table != cur_table) {
      cur_table = h.table;
      const float *tb = a.tables + (size_t)h.table * 46u + 4u * g + r;
#pragma unroll
      for (int o = 0; o < 19; o++) st_raw[o] = tb[o];
      if (PREC != PREC_F32) {
        float ms = st_raw[0];
#pragma unroll
        for (int o = 1; o < 19; o++) ms = fmax_sel(ms, st_raw[o]);
        max_step = wave_max(ms);
      }
    }
    // FP16: scale the chunk so |dequantised| <= (max|level| + 0.2) * max step
    // maps below 2^13; small levels stay far from the FP16 subnormal range
    float s = 1.0f;
    if (PREC != PREC_F32) {
      s2 mx = {0, 0}, mn = {0, 0};
#pragma unroll
      for (int e = 0; e < 4; e++)
#pragma unroll
        for (int hh = 0; hh < 2; hh++) {
          const uint32_t ww[4] = {q[e][hh].x, q[e][hh].y, q[e][hh].z, q[e][hh].w};
#pragma unroll
          for (int i = 0; i < 4; i++) {
            const s2 v = __builtin_bit_cast(s2, ww[i]);
            mx = __builtin_elementwise_max(mx, v);
            mn = __builtin_elementwise_min(mn, v);
          }
        }
      const float ml = fmax_sel(fmax_sel((float)mx[0], (float)mx[1]), fmax_sel(-(float)mn[0], -(float)mn[1]));
      s = pow2_scale((wave_max(ml) + 0.2f) * max_step, SCALE_TOP);
    }
    float st[19];
#pragma unroll
    for (int o = 0; o < 19; o++) st[o] = st_raw[o] * s;
    // passes 1 and 2, one x-frequency tile at a time
    f4 d2[16];
#pragma unroll
    for (int x = 0; x < 16; x++) {
      float v[4];
#pragma unroll
      for (int e = 0; e < 4; e++) {
        const float lf = lane_level(q, e, x);
        v[e] = __builtin_fmaf(fclamp_sel(lf, -1.0f, 1.0f), 0.2f, lf) * st[e + x];
      }
      h4 vh = {}, vl = {};
      if (PREC == PREC_F16) vh = to_h4(v);
      if (PREC == PREC_F16X3) split_h4(v, &vh, &vl);
      const f4 d1 = mma<PREC>(v, cb, vh, vl, bh, bl);
      const float d1v[4] = {d1[0], d1[1], d1[2], d1[3]};
      h4 dh = {}, dl = {};
      if (PREC == PREC_F16) dh = to_h4(d1v);
      if (PREC == PREC_F16X3) split_h4(d1v, &dh, &dl);
      d2[x] = mma<PREC>(cb, d1v, bh, bl, dh, dl);
    }
    const float inv = h.vspan / 255.0f;
    const float scale = inv / s;
    const float bias = __builtin_fmaf(dc, inv, h.vmin);
    uint64_t c0 = 0, nc = 0;
    if (a.corr_off) corr_range(a, c, &c0, &nc);
    if (MFMA3) {
      // pass 3: D3 = C^T (A: row x', k = u) * D2 (B: k = u, column y') per z',
      // B read back from LDS; lane (r, g) slot e gets voxel (z', y' = r, x' = 4g + e)
      uint32_t *tw = lds_w + w * WAVE_WORDS;
#pragma unroll
      for (int e = 0; e < 4; e++)
#pragma unroll
        for (int m = 0; m < 4; m++) {
          const float hv[4] = {d2[4 * m][e], d2[4 * m + 1][e], d2[4 * m + 2][e], d2[4 * m + 3][e]};
          *(h4 *)(tw + (4u * g + e) * P3_SY + r * P3_SZ + 2u * m) = to_h4(hv);
        }
      wave_sync();
#pragma unroll 4
      for (uint32_t zt = 0; zt < 16u; zt++) {
        const h4 b3 = *(const h4 *)(tw + r * P3_SY + zt * P3_SZ + 2u * g);
        const f4 d3 = __builtin_amdgcn_mfma_f32_16x16x16f16(bh, b3, (f4){0.0f, 0.0f, 0.0f, 0.0f}, 0, 0, 0);
        float o[4] = {d3[0], d3[1], d3[2], d3[3]};
        for (uint64_t i = 0; i < nc; i++) {  // corrections in stream order
          const uint32_t pos = __builtin_amdgcn_readfirstlane(a.corr_pos[c0 + i]);
          if (pos >= 4096u) break;
          if ((pos >> 8) != zt) continue;
          const float dv = a.corr_delta[c0 + i] * s;
          const float add = r == ((pos >> 4) & 15u) && g == ((pos & 15u) >> 2) ? dv : 0.0f;
          switch (pos & 3u) {
            case 0: o[0] += add; break;
            case 1: o[1] += add; break;
            case 2: o[2] += add; break;
            default: o[3] += add; break;
          }
        }
        if (F32OUT) {
          *(float4 *)(fdst + zt * 256u + r * 16u + 4u * g) =
              make_float4(__builtin_fmaf(o[0], scale, bias), __builtin_fmaf(o[1], scale, bias),
                          __builtin_fmaf(o[2], scale, bias), __builtin_fmaf(o[3], scale, bias));
        } else {
          uint32_t wd = 0;
#pragma unroll
          for (int e = 0; e < 4; e++) {
            float t = __builtin_fmaf(o[e], scale, bias + 0.5f);  // round half up, clamp, NaN -> 0
            t = fclamp_sel(t, 0.0f, 255.0f);
            wd |= (uint32_t)t << (e * 8);
          }
          uint8_t *p = dst + zt * a.sz + r * a.sy + 4u * g;
          if (((uintptr_t)p & 3u) == 0) {
            *(uint32_t *)p = wd;
          } else {
#pragma unroll
            for (int e = 0; e < 4; e++) p[e] = (uint8_t)(wd >> (8 * e));
          }
        }
      }
      wave_sync();
    } else {
      // pass 3 in registers, corrections and output, one row group
      // (z' = r, y' = 4g + e) at a time
#pragma unroll
      for (int e = 0; e < 4; e++) {
        float in[16], o[16];
#pragma unroll
        for (int x = 0; x < 16; x++) in[x] = d2[x][e];
        idct16_fma(in, o);
        for (uint64_t i = 0; i < nc; i++) {  // corrections in stream order
          const uint32_t pos = __builtin_amdgcn_readfirstlane(a.corr_pos[c0 + i]);
          if (pos >= 4096u) break;
          const uint32_t y = (pos >> 4) & 15u;
          if ((y & 3u) != (uint32_t)e) continue;
          const float dv = a.corr_delta[c0 + i] * s;
          add_at(o, pos & 15u, r == (pos >> 8) && g == (y >> 2) ? dv : 0.0f);
        }
        if (F32OUT) {
          float4 *fd = (float4 *)(fdst + r * 256u + (4u * g + e) * 16u);
#pragma unroll
          for (int i = 0; i < 4; i++)
            fd[i] = make_float4(__builtin_fmaf(o[4 * i], scale, bias), __builtin_fmaf(o[4 * i + 1], scale, bias),
                                __builtin_fmaf(o[4 * i + 2], scale, bias), __builtin_fmaf(o[4 * i + 3], scale, bias));
        } else {
          uint32_t ww[4] = {0, 0, 0, 0};
#pragma unroll
          for (int x = 0; x < 16; x++) {
            float t = __builtin_fmaf(o[x], scale, bias + 0.5f);  // round half up, clamp, NaN -> 0
            t = fclamp_sel(t, 0.0f, 255.0f);
            ww[x >> 2] |= (uint32_t)t << ((x & 3) * 8);
          }
          store_row16(dst + r * a.sz + (4u * g + e) * a.sy, ww);
        }
      }
    }
    if (a.status && lane == 0) a.status[c] = 1;
  }
}

static uint32_t grid_for(const r3d_hip_ctx *ctx, uint32_t n, uint32_t per_cu) {
  uint64_t cap = (uint64_t)(ctx->num_cus > 0 ? ctx->num_cus : 256) * per_cu;
  return (uint32_t)(n < cap ? n : cap);
}

}  // namespace r3d

using namespace r3d;

static int decode_impl(r3d_hip_ctx *ctx, const r3d_chunk_batch *b, const r3d_decode_opts *opts, void *stream,
                       const RegionDst *rg) {
  if (!ctx) return R3D_ERR_INVALID;
  if (!b) return r3d_fail(ctx, R3D_ERR_INVALID, "decode: batch is NULL");
  if (b->nchunks == 0) return R3D_OK;
  if (!b->headers || !b->step_tables || !b->out || b->ntables == 0)
    return r3d_fail(ctx, R3D_ERR_INVALID, "decode: headers, step_tables, out and ntables are required");
  if (b->coef_format == R3D_COEF_DENSE_I16) {
    if (!b->levels) return r3d_fail(ctx, R3D_ERR_INVALID, "decode: dense batch without levels");
    if (((uintptr_t)b->levels & 15u) != 0)
      return r3d_fail(ctx, R3D_ERR_INVALID, "decode: levels must be 16-byte aligned");
  } else if (b->coef_format == R3D_COEF_SPARSE) {
    if (!b->pair_offsets || !b->pairs)
      return r3d_fail(ctx, R3D_ERR_INVALID, "decode: sparse batch without pair_offsets/pairs");
  } else {
    return r3d_fail(ctx, R3D_ERR_INVALID, "decode: unknown coef_format %u", b->coef_format);
  }
  if (b->corr_offsets && (!b->corr_pos || !b->corr_delta))
    return r3d_fail(ctx, R3D_ERR_INVALID, "decode: corr_offsets without corr_pos/corr_delta");
  const uint32_t sy = b->out_stride_y ? b->out_stride_y : 16u;
  const uint32_t sz = b->out_stride_z ? b->out_stride_z : 256u;
  if (sy < 16u || sz < 16u * sy)
    return r3d_fail(ctx, R3D_ERR_INVALID, "decode: strides (%u, %u) overlap rows", sy, sz);
  const uint32_t variant = opts ? opts->variant : R3D_DECODE_EXACT_F32;
  if (variant > R3D_DECODE_MFMA_F16X3)
    return r3d_fail(ctx, R3D_ERR_UNSUPPORTED, "decode: unknown variant %u", variant);
  const bool f32out = opts && (opts->flags & R3D_DECODE_F32_OUT);
  if (opts && (opts->flags & ~R3D_DECODE_F32_OUT))
    return r3d_fail(ctx, R3D_ERR_INVALID, "decode: unknown flags 0x%x", opts->flags);
  if (f32out && (b->out_offsets || b->out_stride_y || b->out_stride_z || ((uintptr_t)b->out & 15u)))
    return r3d_fail(ctx, R3D_ERR_INVALID, "decode: FP32 output is flat and 16-byte aligned");
  if (rg && (variant != R3D_DECODE_EXACT_F32 || f32out || b->out_offsets))
    return r3d_fail(ctx, R3D_ERR_UNSUPPORTED, "decode: a region destination takes the exact u8 decode");
  R3D_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  DecodeArgs a{};
  a.nchunks = b->nchunks;
  a.fmt = b->coef_format;
  a.levels = b->levels;
  a.pair_off = b->pair_offsets;
  a.pairs = b->pairs;
  a.npairs = b->npairs;
  a.hdr = b->headers;
  a.tables = b->step_tables;
  a.ntables = b->ntables;
  a.corr_off = b->corr_offsets;
  a.corr_pos = b->corr_pos;
  a.corr_delta = b->corr_delta;
  a.ncorr = b->ncorr;
  a.out = b->out;
  a.out_off = b->out_offsets;
  a.sy = sy;
  a.sz = sz;
  a.status = b->status;
  a.pair_cnt = b->pair_counts;
  a.corr_cnt = b->corr_counts;
  if (rg) {
    a.rgn = 1;
    a.ncy = rg->ncy;
    a.ncx = rg->ncx;
    a.rfirst = rg->first;
    a.rz = rg->rz;
    a.ry = rg->ry;
    a.rx = rg->rx;
    a.hz = rg->hz;
    a.hy = rg->hy;
    a.hx = rg->hx;
    a.rsy = rg->sy;
    a.rsz = rg->sz;
  }
  hipStream_t st = (hipStream_t)stream;
  if (variant == R3D_DECODE_EXACT_F32) {
    // At most 56 workgroups per CU (8 rounds of the 7 resident), each walking
    // its chunks with a grid stride: the 1024^3 batch (262,144 chunks)
    // 1.35-1.37 -> 1.26-1.28 ms and the 32,768-chunk sparse batch 0.145 ->
    // 0.141 ms against one workgroup per chunk (same bytes). 7 per CU, exactly
    // the resident count, was slower (1.43 ms): with shares that large the
    // CUs do not finish together.
    const uint32_t grid = grid_for(ctx, b->nchunks, 56u);
    const bool sp = b->coef_format == R3D_COEF_SPARSE;
    if (f32out && sp) hipLaunchKernelGGL((k_decode_exact<true, true>), dim3(grid), dim3(256), 0, st, a);
    else if (f32out) hipLaunchKernelGGL((k_decode_exact<true, false>), dim3(grid), dim3(256), 0, st, a);
    else if (sp) hipLaunchKernelGGL((k_decode_exact<false, true>), dim3(grid), dim3(256), 0, st, a);
    else hipLaunchKernelGGL((k_decode_exact<false, false>), dim3(grid), dim3(256), 0, st, a);
  } else {
    // four chunks (waves) per workgroup; 3 waves per SIMD are resident
    const uint32_t grid = grid_for(ctx, (b->nchunks + 3u) / 4u, 8u);
    const bool sparse = b->coef_format == R3D_COEF_SPARSE;
    const int k = (int)(variant - 1u) * 4 + (sparse ? 2 : 0) + (f32out ? 1 : 0);
    static void (*const kernels[12])(DecodeArgs) = {
        k_decode_mfma<PREC_F32, false, false>,   k_decode_mfma<PREC_F32, false, true>,
        k_decode_mfma<PREC_F32, true, false>,    k_decode_mfma<PREC_F32, true, true>,
        k_decode_mfma<PREC_F16, false, false>,   k_decode_mfma<PREC_F16, false, true>,
        k_decode_mfma<PREC_F16, true, false>,    k_decode_mfma<PREC_F16, true, true>,
        k_decode_mfma<PREC_F16X3, false, false>, k_decode_mfma<PREC_F16X3, false, true>,
        k_decode_mfma<PREC_F16X3, true, false>,  k_decode_mfma<PREC_F16X3, true, true>};
    hipLaunchKernelGGL(kernels[k], dim3(grid), dim3(256), 0, st, a);
  }
  R3D_HIP_CHECK(ctx, hipGetLastError());
  return R3D_OK;
}

extern "C" int r3d_hip_decode_chunks(r3d_hip_ctx *ctx, const r3d_chunk_batch *b,
                                     const r3d_decode_opts *opts, void *stream) {
  return decode_impl(ctx, b, opts, stream, nullptr);
}

int r3d::decode_region(r3d_hip_ctx *ctx, const r3d_chunk_batch *b, const RegionDst &rg, hipStream_t stream) {
  return decode_impl(ctx, b, nullptr, (void *)stream, &rg);
}

extern "C" int r3d_hip_encode_chunks(r3d_hip_ctx *ctx, const uint8_t *vol, uint32_t nx, uint32_t ny,
                                     uint32_t nz, const float *step_table, int16_t *levels,
                                     r3d_chunk_header *headers, void *stream) {
  if (!ctx) return R3D_ERR_INVALID;
  if (!vol || !step_table || !levels || !headers)
    return r3d_fail(ctx, R3D_ERR_INVALID, "encode: NULL argument");
  if (nx == 0 || ny == 0 || nz == 0 || nx % 16 || ny % 16 || nz % 16)
    return r3d_fail(ctx, R3D_ERR_INVALID, "encode: dims %ux%ux%u must be positive multiples of 16", nx,
                    ny, nz);
  if (((uintptr_t)vol & 15u) || ((uintptr_t)levels & 15u))
    return r3d_fail(ctx, R3D_ERR_INVALID, "encode: vol and levels must be 16-byte aligned");
  const uint64_t nch = (uint64_t)(nx / 16) * (ny / 16) * (nz / 16);
  if (nch > 0xffffffffull) return r3d_fail(ctx, R3D_ERR_RANGE, "encode: too many chunks");
  R3D_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  EncodeArgs a{vol, nx, ny, nz, nx / 16, ny / 16, (uint32_t)nch, step_table, levels, headers};
  hipLaunchKernelGGL(k_encode, dim3(grid_for(ctx, (uint32_t)nch, 64)), dim3(256), 0,
                     (hipStream_t)stream, a);
  R3D_HIP_CHECK(ctx, hipGetLastError());
  return R3D_OK;
}
