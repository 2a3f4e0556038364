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
// Layout: one 256-lane workgroup per chunk (grid-strided), lane t owns line t
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
};

// dct3d.c:149-154
__device__ __forceinline__ float deq_one(float lv, float step) {
  if (lv == 0.0f) return 0.0f;
  float a = __builtin_fabsf(lv);
  float r = (a - 1.0f) * step + 0.80f * step + 0.40f * step;
  return lv < 0.0f ? -r : r;
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

// dct3d.c:61-68 on a register-resident line. Every k term is added: a zero
// input contributes an exact +-0, so the sums equal the reference's, which
// skips zero inputs (the only possible difference is the sign of an exact
// zero, which the + dc that follows removes).
__device__ __forceinline__ void idct16(const float in[16], float out[16]) {
  float e[8], o[8];
#pragma unroll
  for (int n = 0; n < 8; n++) {
    e[n] = 0.0f;
    o[n] = 0.0f;
  }
#pragma unroll
  for (int k = 0; k < 16; k += 2) {
    const float v = in[k];
#pragma unroll
    for (int n = 0; n < 8; n++) {
      const float p = c_mag[basis_j(k, n)] * v;
      e[n] = basis_neg(k, n) ? e[n] - p : e[n] + p;
    }
  }
#pragma unroll
  for (int k = 1; k < 16; k += 2) {
    const float v = in[k];
#pragma unroll
    for (int n = 0; n < 8; n++) {
      const float p = c_mag[basis_j(k, n)] * v;
      o[n] = basis_neg(k, n) ? o[n] - p : o[n] + p;
    }
  }
#pragma unroll
  for (int n = 0; n < 8; n++) {
    out[n] = e[n] + o[n];
    out[15 - n] = e[n] - o[n];
  }
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

// Three line passes with the transposing LDS write between them. On entry v
// holds line t of the dequantised coefficients; on exit tile holds norm in
// natural (z*16+y)*16+x order.
__device__ __forceinline__ void idct3_tile(float v[16], float *tile, uint32_t t, float dc) {
  float o[16];
  idct16(v, o);
#pragma unroll
  for (int k = 0; k < 16; k++) tile[k * 256 + t] = o[k];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 16; k++) v[k] = tile[t * 16 + k];
  __syncthreads();
  idct16(v, o);
#pragma unroll
  for (int k = 0; k < 16; k++) tile[k * 256 + t] = o[k];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 16; k++) v[k] = tile[t * 16 + k];
  __syncthreads();
  idct16(v, o);
#pragma unroll
  for (int k = 0; k < 16; k++) tile[k * 256 + t] = o[k] + dc;
  __syncthreads();
}

__global__ __launch_bounds__(256) void k_decode_exact(DecodeArgs a) {
  __shared__ float tile[4096];
  __shared__ float steps[48];
  const uint32_t t = threadIdx.x;
  const uint32_t z = t >> 4, y = t & 15u;
  for (uint32_t c = blockIdx.x; c < a.nchunks; c += gridDim.x) {
    const r3d_chunk_header h = a.hdr[c];
    const float dc = (float)h.dcq / 16.0f;
    const bool ok = h.table < a.ntables && h.vspan > 0.0f && __builtin_isfinite(h.vmin) &&
                    __builtin_isfinite(dc);
    uint8_t *dst = a.out + (a.out_off ? a.out_off[c] : (uint64_t)c * 4096u) + z * a.sz + y * a.sy;
    if (!ok) {  // dct3d.c:555-556: rejected header; the chunk is zero-filled
      const uint32_t zw[4] = {0, 0, 0, 0};
      store_row16(dst, zw);
      if (a.status && t == 0) a.status[c] = 0;
      continue;
    }
    if (t < 46) steps[t] = a.tables[(size_t)h.table * 46u + t];
    float v[16];
    if (a.fmt == R3D_COEF_DENSE_I16) {
      const int4 *src = (const int4 *)(a.levels + (size_t)c * 4096u + t * 16u);
      const int4 q0 = src[0], q1 = src[1];
      const int32_t w[8] = {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, q1.z, q1.w};
      __syncthreads();
#pragma unroll
      for (int k = 0; k < 16; k++) {
        const int16_t l = (int16_t)((uint32_t)w[k >> 1] >> ((k & 1) * 16));
        v[k] = deq_one((float)l, steps[z + y + k]);
      }
    } else {
#pragma unroll
      for (int k = 0; k < 16; k++) tile[k * 256 + t] = 0.0f;
      __syncthreads();
      // offsets are caller data: clamp so a corrupt batch cannot read past pairs
      const uint64_t p1 = a.pair_off[c + 1] < a.npairs ? a.pair_off[c + 1] : a.npairs;
      for (uint64_t p = (uint64_t)a.pair_off[c] + t; p < p1; p += 256u) {
        const uint32_t pr = a.pairs[p];
        const uint32_t pos = pr & 4095u;
        const float l = (float)(int16_t)(pr >> 16);
        tile[pos] = deq_one(l, steps[(pos >> 8) + ((pos >> 4) & 15u) + (pos & 15u)]);
      }
      __syncthreads();
#pragma unroll
      for (int k = 0; k < 16; k++) v[k] = tile[t * 16 + k];
      __syncthreads();
    }
    idct3_tile(v, tile, t, dc);
    if (a.corr_off) {  // dct3d.c:572-587: serial, in stream order
      if (t == 0) {
        const uint64_t c0 = a.corr_off[c];
        const uint64_t c1 = a.corr_off[c + 1] < a.ncorr ? a.corr_off[c + 1] : a.ncorr;
        const uint64_t n = c1 > c0 ? (c1 - c0 > 4096u ? 4096u : c1 - c0) : 0;
        for (uint64_t i = 0; i < n; i++) {
          const uint32_t pos = a.corr_pos[c0 + i];
          if (pos >= 4096u) break;
          tile[pos] += a.corr_delta[c0 + i];
        }
      }
      __syncthreads();
    }
    const float inv = h.vspan / 255.0f;
    uint32_t w[4] = {0, 0, 0, 0};
#pragma unroll
    for (int x = 0; x < 16; x++) {
      const float val = tile[t * 16 + x] * inv + h.vmin;
      float r = __builtin_roundf(val);
      r = !(r >= 0.0f) ? 0.0f : (r > 255.0f ? 255.0f : r);
      w[x >> 2] |= (uint32_t)r << ((x & 3) * 8);
    }
    store_row16(dst, w);
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

static uint32_t grid_for(const r3d_hip_ctx *ctx, uint32_t n, uint32_t per_cu) {
  uint64_t cap = (uint64_t)(ctx->num_cus > 0 ? ctx->num_cus : 256) * per_cu;
  return (uint32_t)(n < cap ? n : cap);
}

}  // namespace r3d

using namespace r3d;

extern "C" int r3d_hip_decode_chunks(r3d_hip_ctx *ctx, const r3d_chunk_batch *b,
                                     const r3d_decode_opts *opts, void *stream) {
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
  if (variant != R3D_DECODE_EXACT_F32)
    return r3d_fail(ctx, R3D_ERR_UNSUPPORTED, "decode: variant %u not built", variant);
  R3D_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  DecodeArgs a;
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
  const uint32_t grid = grid_for(ctx, b->nchunks, 1u << 16);
  hipLaunchKernelGGL(k_decode_exact, dim3(grid), dim3(256), 0, (hipStream_t)stream, a);
  R3D_HIP_CHECK(ctx, hipGetLastError());
  return R3D_OK;
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
