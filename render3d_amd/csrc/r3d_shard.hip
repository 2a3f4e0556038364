// r3d_shard.hip -- zarr v3 shards of 3ddct chunks on gfx950: the index
// crc32c, the index lookups and the region decode of r3d_shard_decode_region
// (src/shard/shardio.c:39-306).
//
// A region decode is, per slab of at most 262,144 chunks of the region's
// chunk box: k_shard_gather (index entry -> blob range, or missing / corrupt),
// k_parse_blobs (the entropy stage), the exact decode writing straight into
// the region with clipped row stores, and k_shard_finish (per-chunk status
// and counts). Each listed shard's index crc is checked first by
// k_crc_pieces + k_crc_combine. All of it is integer or byte work bound by
// the parse (one lane per blob); the crc and the index lookups read 4 MiB per
// shard.
#include <string.h>

#include "r3d_internal.h"

namespace r3d {

constexpr uint32_t SH_GRID = 64u;
constexpr uint32_t SH_N = SH_GRID * SH_GRID * SH_GRID;
constexpr uint64_t SH_INDEX = (uint64_t)SH_N * 16u + 4u;
constexpr uint32_t SH_SLAB = SH_N;          // chunks per gather/parse/decode pass
constexpr uint32_t SH_CAP = 4096u;          // levels and corrections per chunk
constexpr uint32_t CRC_PIECE = 1024u;       // bytes per lane of k_crc_pieces
constexpr uint32_t CRC_NT = 256u;
constexpr uint32_t CRC_COMBINE_NT = 1024u;
constexpr uint32_t CRC_POLY = 0x82F63B78u;  // Castagnoli, reflected

__constant__ uint32_t c_crc_tab[256];  // byte-at-a-time table of the reflected polynomial
__constant__ uint32_t c_x2k[64];       // x^(2^k) mod P in the reflected representation

// Products in GF(2)[x] / P, reflected: bit 31 holds the coefficient of x^0.
// With them crc(A B) = crc(A) * x^(8 |B|) + crc(B) for the finished CRC
// (initial value and final xor all ones), so pieces are checked independently
// and folded in order.
__device__ uint32_t gf_mul(uint32_t a, uint32_t b) {
  uint32_t r = 0;
  for (int j = 0; j < 32; j++) {
    r ^= (a & (0x80000000u >> j)) ? b : 0u;
    b = (b >> 1) ^ (CRC_POLY & (0u - (b & 1u)));
  }
  return r;
}

__device__ uint32_t gf_x8n(uint64_t n) {  // x^(8 n) mod P, n < 2^61
  uint32_t r = 0x80000000u;
  uint64_t e = n << 3;
  for (int k = 0; e; k++, e >>= 1)
    if (e & 1u) r = gf_mul(r, c_x2k[k]);
  return r;
}

// CRC of each CRC_PIECE-byte piece of data[0, n). The pieces are read as the
// aligned dwords covering them: an aligned dword holding a byte of the data
// never crosses a page the data does not touch.
__global__ __launch_bounds__(CRC_NT) void k_crc_pieces(const uint8_t *data, uint64_t n, uint32_t npieces,
                                                       uint32_t *part) {
  __shared__ uint32_t tab[256];
  tab[threadIdx.x] = c_crc_tab[threadIdx.x];
  __syncthreads();
  const uint32_t i = blockIdx.x * CRC_NT + threadIdx.x;
  if (i >= npieces) return;
  const uint64_t b0 = (uint64_t)i * CRC_PIECE;
  const uint64_t b1 = n - b0 < CRC_PIECE ? n : b0 + CRC_PIECE;
  const uintptr_t a0 = (uintptr_t)(data + b0), a1 = (uintptr_t)(data + b1);
  uint32_t c = 0xFFFFFFFFu;
  for (uintptr_t a = a0 & ~(uintptr_t)3u; a < a1; a += 4) {
    const uint32_t w = *(const uint32_t *)a;
#pragma unroll
    for (uint32_t k = 0; k < 4; k++)
      if (a + k >= a0 && a + k < a1) c = tab[(c ^ (w >> (8u * k))) & 0xFFu] ^ (c >> 8);
  }
  part[i] = ~c;
}

// Folds the pieces' CRCs in order (one workgroup): each lane folds a run of
// consecutive pieces, then a tree. Writes the CRC to *out and, when want is
// given, whether it equals the little-endian u32 at want to *ok.
__global__ __launch_bounds__(CRC_COMBINE_NT) void k_crc_combine(const uint32_t *part, uint32_t npieces, uint64_t n,
                                                                uint32_t *out, const uint8_t *want, uint32_t *ok) {
  __shared__ uint32_t s_crc[CRC_COMBINE_NT];
  __shared__ uint64_t s_len[CRC_COMBINE_NT];
  const uint32_t t = threadIdx.x;
  const uint32_t q = (npieces + CRC_COMBINE_NT - 1u) / CRC_COMBINE_NT;
  const uint32_t p0 = t * q < npieces ? t * q : npieces, p1 = p0 + q < npieces ? p0 + q : npieces;
  const uint32_t xp = gf_x8n(CRC_PIECE);
  uint32_t crc = 0;
  uint64_t len = 0;
  for (uint32_t i = p0; i < p1; i++) {
    const uint64_t li = n - (uint64_t)i * CRC_PIECE < CRC_PIECE ? n - (uint64_t)i * CRC_PIECE : CRC_PIECE;
    crc = gf_mul(crc, li == CRC_PIECE ? xp : gf_x8n(li)) ^ part[i];
    len += li;
  }
  s_crc[t] = crc;
  s_len[t] = len;
  __syncthreads();
  for (uint32_t h = 1; h < CRC_COMBINE_NT; h <<= 1) {
    if ((t & (2u * h - 1u)) == 0u) {
      const uint32_t b = s_crc[t + h];
      const uint64_t lb = s_len[t + h];
      s_crc[t] = gf_mul(s_crc[t], gf_x8n(lb)) ^ b;
      s_len[t] += lb;
    }
    __syncthreads();
  }
  if (t == 0) {
    const uint32_t v = s_crc[0];
    if (out) *out = v;
    if (want) {
      const uint32_t w = (uint32_t)want[0] | (uint32_t)want[1] << 8 | (uint32_t)want[2] << 16 |
                         (uint32_t)want[3] << 24;
      *ok = v == w ? 1u : 0u;
    }
  }
}

// Device view of the shard box: one entry per shard coordinate in the box
// (z-major); absent shards have present 0.
struct ShardDev {
  const uint8_t *bytes;
  uint64_t nbytes;
  uint32_t present, ok;  // ok: long enough for the index and crc matched
};

struct GatherArgs {
  uint32_t nch, first;         // chunks of this pass; box cell of its chunk 0
  uint32_t ncy, ncx;           // chunk box dims (y, x)
  uint64_t cz0, cy0, cx0;      // chunk box origin, global chunk coordinates
  uint64_t sz0, sy0, sx0;      // shard box origin
  uint32_t nsy, nsx;           // shard box dims (y, x)
  const ShardDev *shards;
  const uint8_t *base;         // blob offsets are relative to base
  uint64_t *offsets;
  uint32_t *lengths;
  int32_t *state;              // 1 present, 0 missing / absent, -1 corrupt
};

__device__ __forceinline__ uint64_t ld_le64(const uint8_t *p) {
  uint64_t v = 0;
#pragma unroll
  for (int i = 7; i >= 0; i--) v = (v << 8) | p[i];
  return v;
}

// shardio.c:157-172 for every chunk of the pass.
__global__ __launch_bounds__(256) void k_shard_gather(GatherArgs a) {
  const uint32_t c = blockIdx.x * 256u + threadIdx.x;
  if (c >= a.nch) return;
  const uint32_t cl = a.first + c;
  const uint32_t ix = cl % a.ncx, r = cl / a.ncx, iy = r % a.ncy, iz = r / a.ncy;
  const uint64_t gz = a.cz0 + iz, gy = a.cy0 + iy, gx = a.cx0 + ix;
  const ShardDev sh =
      a.shards[((gz / SH_GRID - a.sz0) * a.nsy + (gy / SH_GRID - a.sy0)) * a.nsx + (gx / SH_GRID - a.sx0)];
  int32_t st = 0;
  uint64_t off = 0;
  uint32_t len = 0;
  if (sh.present && !sh.ok) {
    st = -1;
  } else if (sh.present) {
    const uint64_t ci = ((gz % SH_GRID) * SH_GRID + gy % SH_GRID) * SH_GRID + gx % SH_GRID;
    const uint8_t *e = sh.bytes + (sh.nbytes - SH_INDEX) + ci * 16u;
    const uint64_t o = ld_le64(e), l = ld_le64(e + 8);
    const uint64_t payload = sh.nbytes - SH_INDEX;
    if (o == ~0ull && l == ~0ull) {
      st = 0;
    } else if (o == ~0ull || l == ~0ull || o > payload || l > payload - o) {
      st = -1;
    } else {
      st = 1;
      off = (uint64_t)(uintptr_t)(sh.bytes + o) - (uint64_t)(uintptr_t)a.base;
      // a 3ddct stream is read at most a few tens of KB deep; past 4 GiB the
      // parser's zero tail stands in for bytes it never reaches
      len = l > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)l;
    }
  }
  a.offsets[c] = off;
  a.lengths[c] = len;  // 0 for missing and corrupt chunks: the parse rejects them, the decode zero-fills
  a.state[c] = st;
}

// Per-chunk status and the four counters (decoded, missing, corrupt,
// beyond int16 / capacity).
__global__ __launch_bounds__(256) void k_shard_finish(const int32_t *state, const int32_t *parse_st, uint32_t nch,
                                                      int32_t *status, uint32_t *counts) {
  const uint32_t c = blockIdx.x * 256u + threadIdx.x;
  int32_t f = 3;  // lanes past the end count nowhere
  if (c < nch) {
    const int32_t s = state[c], p = parse_st[c];
    f = s == 0 ? 0 : s < 0 ? -1 : p == 1 ? 1 : p == 0 ? -1 : -2;
    if (status) status[c] = f;
  }
  if (counts) {
    const uint64_t m1 = __ballot(f == 1), m0 = __ballot(f == 0), mc = __ballot(f == -1), mw = __ballot(f == -2);
    if ((threadIdx.x & 63u) == 0u) {
      if (m1) atomicAdd(&counts[0], (uint32_t)__popcll(m1));
      if (m0) atomicAdd(&counts[1], (uint32_t)__popcll(m0));
      if (mc) atomicAdd(&counts[2], (uint32_t)__popcll(mc));
      if (mw) atomicAdd(&counts[3], (uint32_t)__popcll(mw));
    }
  }
}

// Zero n bytes (the parts of a region outside the volume).
__global__ __launch_bounds__(256) void k_zero(uint8_t *p, uint64_t n) {
  const uint64_t head = (16u - ((uintptr_t)p & 15u)) & 15u;
  const uint64_t h = head < n ? head : n;
  const uint64_t nv = (n - h) / 16u;
  const uint64_t i0 = (uint64_t)blockIdx.x * 256u + threadIdx.x, step = (uint64_t)gridDim.x * 256u;
  uint4 *v = (uint4 *)(p + h);
  for (uint64_t i = i0; i < nv; i += step) v[i] = make_uint4(0, 0, 0, 0);
  if (i0 < h) p[i0] = 0;
  const uint64_t tail0 = h + nv * 16u;
  if (i0 < n - tail0) p[tail0 + i0] = 0;
}

int upload_crc_tables(r3d_hip_ctx *ctx) {
  uint32_t tab[256], x2k[64];
  for (uint32_t i = 0; i < 256; i++) {
    uint32_t c = i;
    for (int k = 0; k < 8; k++) c = (c >> 1) ^ (CRC_POLY & (0u - (c & 1u)));
    tab[i] = c;
  }
  // x^(2^k) mod P: x^1, then squares (reflected products, as gf_mul)
  auto mul = [](uint32_t a, uint32_t b) {
    uint32_t r = 0;
    for (int j = 0; j < 32; j++) {
      r ^= (a & (0x80000000u >> j)) ? b : 0u;
      b = (b >> 1) ^ (CRC_POLY & (0u - (b & 1u)));
    }
    return r;
  };
  x2k[0] = 0x40000000u;
  for (int k = 1; k < 64; k++) x2k[k] = mul(x2k[k - 1], x2k[k - 1]);
  R3D_HIP_CHECK(ctx, hipMemcpyToSymbol(HIP_SYMBOL(c_crc_tab), tab, sizeof tab, 0, hipMemcpyHostToDevice));
  R3D_HIP_CHECK(ctx, hipMemcpyToSymbol(HIP_SYMBOL(c_x2k), x2k, sizeof x2k, 0, hipMemcpyHostToDevice));
  return R3D_OK;
}

int workspace(r3d_hip_ctx *ctx, size_t bytes, void **out) {
  if (ctx->ws_cap < bytes) {
    void *old = ctx->d_ws;
    ctx->d_ws = nullptr;
    ctx->ws_cap = 0;
    if (old) R3D_HIP_CHECK(ctx, hipFree(old));
    if (hipMalloc(&ctx->d_ws, bytes) != hipSuccess) {
      ctx->d_ws = nullptr;
      return r3d_fail(ctx, R3D_ERR_NOMEM, "workspace of %zu bytes", bytes);
    }
    ctx->ws_cap = bytes;
  }
  *out = ctx->d_ws;
  return R3D_OK;
}

static int crc_launch(r3d_hip_ctx *ctx, const uint8_t *data, uint64_t n, uint32_t *part, uint32_t *out,
                      const uint8_t *want, uint32_t *ok, hipStream_t st) {
  const uint64_t np = (n + CRC_PIECE - 1u) / CRC_PIECE;
  if (np > 0xFFFFFFFFull / CRC_NT) return r3d_fail(ctx, R3D_ERR_RANGE, "crc32c: %llu bytes", (unsigned long long)n);
  if (np) hipLaunchKernelGGL(k_crc_pieces, dim3((uint32_t)((np + CRC_NT - 1u) / CRC_NT)), dim3(CRC_NT), 0, st, data, n,
                             (uint32_t)np, part);
  hipLaunchKernelGGL(k_crc_combine, dim3(1), dim3(CRC_COMBINE_NT), 0, st, part, (uint32_t)np, n, out, want, ok);
  R3D_HIP_CHECK(ctx, hipGetLastError());
  return R3D_OK;
}

// Per-chunk workspace of one pass: gather outputs, parse outputs, decode status.
struct SlabWs {
  uint64_t *offsets;
  uint32_t *lengths;
  int32_t *state, *parse_st;
  r3d_chunk_header *hdr;
  float *steps;
  uint32_t *pair_off, *pair_cnt, *pairs, *corr_off, *corr_cnt;
  uint16_t *corr_pos;
  float *corr_delta;
};

static size_t slab_ws(uint32_t n, uint8_t *base, SlabWs *w) {
  size_t at = 0;
  auto take = [&](size_t bytes) {
    uint8_t *p = base ? base + at : nullptr;
    at += (bytes + 255u) & ~(size_t)255u;
    return p;
  };
  w->offsets = (uint64_t *)take(8ull * n);
  w->lengths = (uint32_t *)take(4ull * n);
  w->state = (int32_t *)take(4ull * n);
  w->parse_st = (int32_t *)take(4ull * n);
  w->hdr = (r3d_chunk_header *)take(sizeof(r3d_chunk_header) * (size_t)n);
  w->steps = (float *)take(46ull * 4u * n);
  w->pair_off = (uint32_t *)take(4ull * (n + 1u));
  w->pair_cnt = (uint32_t *)take(4ull * n);
  w->pairs = (uint32_t *)take(4ull * SH_CAP * n);
  w->corr_off = (uint32_t *)take(4ull * (n + 1u));
  w->corr_cnt = (uint32_t *)take(4ull * n);
  w->corr_pos = (uint16_t *)take(2ull * SH_CAP * n);
  w->corr_delta = (float *)take(4ull * SH_CAP * n);
  return at;
}

}  // namespace r3d

using namespace r3d;

namespace {
struct Box {
  uint64_t z1, y1, x1;          // clamped exclusive bounds (volume voxels)
  uint64_t cz0, cy0, cx0;       // first chunk
  uint64_t ncz, ncy, ncx;       // chunk box dims
  bool empty;
};
Box region_box(const r3d_shard_region *r) {
  Box b{};
  b.empty = r->dz == 0 || r->dy == 0 || r->dx == 0 || r->z0 >= r->nz || r->y0 >= r->ny || r->x0 >= r->nx;
  if (b.empty) return b;
  b.z1 = r->dz < r->nz - r->z0 ? r->z0 + r->dz : r->nz;
  b.y1 = r->dy < r->ny - r->y0 ? r->y0 + r->dy : r->ny;
  b.x1 = r->dx < r->nx - r->x0 ? r->x0 + r->dx : r->nx;
  b.cz0 = r->z0 / 16u;
  b.cy0 = r->y0 / 16u;
  b.cx0 = r->x0 / 16u;
  b.ncz = (b.z1 - 1u) / 16u - b.cz0 + 1u;
  b.ncy = (b.y1 - 1u) / 16u - b.cy0 + 1u;
  b.ncx = (b.x1 - 1u) / 16u - b.cx0 + 1u;
  return b;
}
}  // namespace

extern "C" uint64_t r3d_hip_shard_region_chunks(const r3d_shard_region *region) {
  if (!region) return 0;
  const Box b = region_box(region);
  return b.empty ? 0 : b.ncz * b.ncy * b.ncx;
}

extern "C" int r3d_hip_crc32c(r3d_hip_ctx *ctx, const uint8_t *data, uint64_t n, uint32_t *crc, void *stream) {
  if (!ctx) return R3D_ERR_INVALID;
  if (!crc || (n && !data)) return r3d_fail(ctx, R3D_ERR_INVALID, "crc32c: NULL argument");
  R3D_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  const uint64_t np = (n + CRC_PIECE - 1u) / CRC_PIECE;
  void *ws = nullptr;
  int rc = workspace(ctx, (size_t)(np ? np : 1u) * 4u, &ws);
  if (rc) return rc;
  return crc_launch(ctx, data, n, (uint32_t *)ws, crc, nullptr, nullptr, (hipStream_t)stream);
}

int r3d::shard_decode_bricks(r3d_hip_ctx *ctx, const r3d_shard_src *shards, uint32_t nshards, const uint64_t dims[3],
                             const BrickTarget *t, uint32_t n, uint8_t *atlas, hipStream_t st) {
  const uint64_t ncz = dims[0] / 16u, ncy = dims[1] / 16u, ncx = dims[2] / 16u;
  if (!n) return R3D_OK;
  if (n > 64u || dims[0] % 16u || dims[1] % 16u || dims[2] % 16u)
    return r3d_fail(ctx, R3D_ERR_INVALID, "shard bricks: %u bricks of a store of %llux%llux%llu", n,
                    (unsigned long long)dims[2], (unsigned long long)dims[1], (unsigned long long)dims[0]);
  // each brick's chunk box (8^3, clamped to the store), and where each chunk goes
  struct JobBox {
    uint64_t cz0, cy0, cx0;
    uint32_t dz, dy, dx, first;
  };
  std::vector<JobBox> jb(n);
  std::vector<uint64_t> outoff;
  uint32_t N = 0;
  for (uint32_t j = 0; j < n; j++) {
    JobBox &b = jb[j];
    b.cz0 = (uint64_t)t[j].bz * 8u, b.cy0 = (uint64_t)t[j].by * 8u, b.cx0 = (uint64_t)t[j].bx * 8u;
    b.dz = b.cz0 < ncz ? (uint32_t)(ncz - b.cz0 < 8u ? ncz - b.cz0 : 8u) : 0u;
    b.dy = b.cy0 < ncy ? (uint32_t)(ncy - b.cy0 < 8u ? ncy - b.cy0 : 8u) : 0u;
    b.dx = b.cx0 < ncx ? (uint32_t)(ncx - b.cx0 < 8u ? ncx - b.cx0 : 8u) : 0u;
    b.first = N;
    const uint64_t slot = (uint64_t)t[j].slot * 128u * 128u * 128u;
    for (uint32_t iz = 0; iz < b.dz; iz++)
      for (uint32_t iy = 0; iy < b.dy; iy++)
        for (uint32_t ix = 0; ix < b.dx; ix++)
          outoff.push_back(slot + (uint64_t)iz * 16u * 16384u + (uint64_t)iy * 16u * 128u + (uint64_t)ix * 16u);
    N += b.dz * b.dy * b.dx;
  }
  // the store's shard grid and its table; a shard listed twice: the last one counts
  const uint64_t nsz = (ncz + SH_GRID - 1u) / SH_GRID, nsy = (ncy + SH_GRID - 1u) / SH_GRID,
                 nsx = (ncx + SH_GRID - 1u) / SH_GRID;
  const uint64_t nsb = nsz * nsy * nsx;
  std::vector<ShardDev> tab(nsb);
  const uint8_t *base = nullptr;
  for (uint32_t k = 0; k < nshards; k++) {
    const r3d_shard_src &q = shards[k];
    if (q.sz >= nsz || q.sy >= nsy || q.sx >= nsx) continue;
    ShardDev &d = tab[((uint64_t)q.sz * nsy + q.sy) * nsx + q.sx];
    d.bytes = q.bytes;
    d.nbytes = q.nbytes;
    d.present = 1;
    d.ok = 0;
    if (q.nbytes >= SH_INDEX && (!base || q.bytes < base)) base = q.bytes;
  }
  if (!base) base = atlas;  // no readable shard: every length is 0, nothing is read
  SlabWs w;
  const size_t slab_bytes = (slab_ws(N ? N : 1u, nullptr, &w) + 255u) & ~(size_t)255u;
  const size_t crc_bytes = (size_t)((SH_N * 16ull + CRC_PIECE - 1u) / CRC_PIECE) * 4u;
  void *ws = nullptr;
  int rc = workspace(ctx, slab_bytes + crc_bytes, &ws);
  if (rc) return rc;
  slab_ws(N ? N : 1u, (uint8_t *)ws, &w);
  uint32_t *crc_part = (uint32_t *)((uint8_t *)ws + slab_bytes);
  // one staging slot: [shard table][chunk destinations]
  const size_t tb = ((sizeof(ShardDev) * nsb + 15u) & ~(size_t)15u);
  std::vector<uint8_t> buf(tb + 8u * (size_t)N);
  memcpy(buf.data(), tab.data(), sizeof(ShardDev) * nsb);
  if (N) memcpy(buf.data() + tb, outoff.data(), 8u * (size_t)N);
  ViewSlot *sl = nullptr;
  rc = stage(ctx, buf.data(), buf.size(), st, &sl);
  if (rc) return rc;
  ShardDev *dtab = (ShardDev *)sl->dev;
  const uint64_t *doff = (const uint64_t *)((const uint8_t *)sl->dev + tb);
  for (uint64_t i = 0; i < nsb && rc == R3D_OK; i++) {
    if (!tab[i].present || tab[i].nbytes < SH_INDEX) continue;
    const uint8_t *index = tab[i].bytes + (tab[i].nbytes - SH_INDEX);
    rc = crc_launch(ctx, index, (uint64_t)SH_N * 16u, crc_part, nullptr, index + (uint64_t)SH_N * 16u, &dtab[i].ok,
                    st);
  }
  for (uint32_t j = 0; j < n && rc == R3D_OK; j++) {  // bricks reaching past the store: zeros first
    const JobBox &b = jb[j];
    if (b.dz < 8u || b.dy < 8u || b.dx < 8u)
      hipLaunchKernelGGL(k_zero, dim3(512), dim3(256), 0, st, atlas + (uint64_t)t[j].slot * 128u * 128u * 128u,
                         (uint64_t)128u * 128u * 128u);
    const uint32_t nch = b.dz * b.dy * b.dx;
    if (!nch) continue;
    GatherArgs g{nch,   0u, b.dy, b.dx,  b.cz0, b.cy0, b.cx0, 0u,        0u,         0u,
                 (uint32_t)nsy, (uint32_t)nsx, dtab, base, w.offsets + b.first, w.lengths + b.first,
                 w.state + b.first};
    hipLaunchKernelGGL(k_shard_gather, dim3((nch + 255u) / 256u), dim3(256), 0, st, g);
  }
  if (rc == R3D_OK && hipGetLastError() != hipSuccess) rc = r3d_fail(ctx, R3D_ERR_DEVICE, "shard bricks: launch");
  if (rc == R3D_OK && N) {
    r3d_blob_batch in{N, 1u, SH_CAP, 0u, base, w.offsets, w.lengths};
    r3d_blob_parse_out po{w.hdr,     w.steps,    w.pair_off, w.pair_cnt,  w.pairs,
                          w.corr_off, w.corr_cnt, w.corr_pos, w.corr_delta, w.parse_st};
    rc = r3d_hip_parse_blobs(ctx, &in, &po, (void *)st);
    if (rc == R3D_OK) {
      r3d_chunk_batch cb{};
      cb.nchunks = N;
      cb.coef_format = R3D_COEF_SPARSE;
      cb.pair_offsets = w.pair_off;
      cb.pairs = w.pairs;
      cb.npairs = (uint64_t)N * SH_CAP;
      cb.headers = w.hdr;
      cb.step_tables = w.steps;
      cb.ntables = N;
      cb.corr_offsets = w.corr_off;
      cb.corr_pos = w.corr_pos;
      cb.corr_delta = w.corr_delta;
      cb.ncorr = (uint64_t)N * SH_CAP;
      cb.out = atlas;
      cb.out_offsets = doff;
      cb.out_stride_y = 128u;
      cb.out_stride_z = 16384u;
      cb.pair_counts = w.pair_cnt;
      cb.corr_counts = w.corr_cnt;
      const r3d_decode_opts o{R3D_DECODE_EXACT_F32, 0u};
      rc = r3d_hip_decode_chunks(ctx, &cb, &o, (void *)st);
    }
  }
  // the staged table and destinations are read up to the decode
  if (hipEventRecord(sl->ev, st) == hipSuccess) sl->recorded = true;
  return rc;
}

extern "C" int r3d_hip_shard_decode_region(r3d_hip_ctx *ctx, const r3d_shard_src *shards, uint32_t nshards,
                                           const r3d_shard_region *region, uint8_t *dst, int32_t *chunk_status,
                                           uint32_t *counts, void *stream) {
  if (!ctx) return R3D_ERR_INVALID;
  if (!region || !dst) return r3d_fail(ctx, R3D_ERR_INVALID, "shard_decode_region: NULL region or dst");
  if (nshards && !shards) return r3d_fail(ctx, R3D_ERR_INVALID, "shard_decode_region: NULL shard list");
  const r3d_shard_region &r = *region;
  if (r.dz == 0 || r.dy == 0 || r.dx == 0)
    return r3d_fail(ctx, R3D_ERR_INVALID, "shard_decode_region: empty region %ux%ux%u", r.dz, r.dy, r.dx);
  const uint64_t plane = (uint64_t)r.dy * r.dx;
  if (plane > (1ull << 40) || (uint64_t)r.dz > (1ull << 40) / plane)  // shardio.c:272-275, and HBM
    return r3d_fail(ctx, R3D_ERR_RANGE, "shard_decode_region: region %ux%ux%u is too large", r.dz, r.dy, r.dx);
  for (uint32_t s = 0; s < nshards; s++)
    if (shards[s].nbytes && !shards[s].bytes)
      return r3d_fail(ctx, R3D_ERR_INVALID, "shard_decode_region: shard %u has no bytes", s);
  const uint64_t total = plane * r.dz;
  const Box b = region_box(&r);
  const uint64_t nbox = b.empty ? 0 : b.ncz * b.ncy * b.ncx;
  if (nbox > 0xFFFFFFFFull) return r3d_fail(ctx, R3D_ERR_RANGE, "shard_decode_region: too many chunks");
  R3D_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  hipStream_t st = (hipStream_t)stream;
  const bool full = !b.empty && b.z1 - r.z0 == r.dz && b.y1 - r.y0 == r.dy && b.x1 - r.x0 == r.dx;
  if (!full) {  // shardio.c:270: zeros wherever no chunk writes
    const uint64_t blocks = (total / 16u + 255u) / 256u;
    const uint32_t grid = (uint32_t)(blocks < 1u ? 1u : blocks > 4096u ? 4096u : blocks);
    hipLaunchKernelGGL(k_zero, dim3(grid), dim3(256), 0, st, dst, total);
    R3D_HIP_CHECK(ctx, hipGetLastError());
  }
  if (b.empty) return R3D_OK;
  // the shard box and its table; a shard listed twice: the last one counts
  const uint64_t sz0 = b.cz0 / SH_GRID, sy0 = b.cy0 / SH_GRID, sx0 = b.cx0 / SH_GRID;
  const uint64_t nsz = (b.cz0 + b.ncz - 1u) / SH_GRID - sz0 + 1u, nsy = (b.cy0 + b.ncy - 1u) / SH_GRID - sy0 + 1u,
                 nsx = (b.cx0 + b.ncx - 1u) / SH_GRID - sx0 + 1u;
  const uint64_t nsb = nsz * nsy * nsx;
  ShardDev *tab = new ShardDev[nsb]();
  const uint8_t *base = nullptr;
  for (uint32_t s = 0; s < nshards; s++) {
    const r3d_shard_src &q = shards[s];
    if (q.sz < sz0 || q.sz >= sz0 + nsz || q.sy < sy0 || q.sy >= sy0 + nsy || q.sx < sx0 || q.sx >= sx0 + nsx)
      continue;
    ShardDev &d = tab[((q.sz - sz0) * nsy + (q.sy - sy0)) * nsx + (q.sx - sx0)];
    d.bytes = q.bytes;
    d.nbytes = q.nbytes;
    d.present = 1;
    d.ok = 0;  // set on the device by the crc check
    if (q.nbytes >= SH_INDEX && (!base || q.bytes < base)) base = q.bytes;
  }
  if (!base) base = dst;  // no readable shard: every length is 0, nothing is read
  const uint32_t slab = nbox < SH_SLAB ? (uint32_t)nbox : SH_SLAB;
  SlabWs w;
  const size_t slab_bytes = slab_ws(slab, nullptr, &w);
  const size_t crc_bytes = (size_t)((SH_N * 16ull + CRC_PIECE - 1u) / CRC_PIECE) * 4u;
  void *ws = nullptr;
  int rc = workspace(ctx, slab_bytes + crc_bytes, &ws);
  if (rc) {
    delete[] tab;
    return rc;
  }
  slab_ws(slab, (uint8_t *)ws, &w);
  uint32_t *crc_part = (uint32_t *)((uint8_t *)ws + slab_bytes);
  ViewSlot *sl = nullptr;
  rc = stage(ctx, tab, sizeof(ShardDev) * nsb, st, &sl);
  if (rc) {
    delete[] tab;
    return rc;
  }
  ShardDev *dtab = (ShardDev *)sl->dev;
  for (uint64_t i = 0; i < nsb && rc == R3D_OK; i++) {
    if (!tab[i].present || tab[i].nbytes < SH_INDEX) continue;  // too short: stays corrupt (ok = 0)
    const uint8_t *index = tab[i].bytes + (tab[i].nbytes - SH_INDEX);
    rc = crc_launch(ctx, index, (uint64_t)SH_N * 16u, crc_part, nullptr, index + (uint64_t)SH_N * 16u, &dtab[i].ok,
                    st);
  }
  delete[] tab;
  RegionDst rg{};
  rg.ncz = (uint32_t)b.ncz;
  rg.ncy = (uint32_t)b.ncy;
  rg.ncx = (uint32_t)b.ncx;
  rg.rz = (int64_t)(b.cz0 * 16u) - (int64_t)r.z0;
  rg.ry = (int64_t)(b.cy0 * 16u) - (int64_t)r.y0;
  rg.rx = (int64_t)(b.cx0 * 16u) - (int64_t)r.x0;
  rg.hz = (int64_t)(b.z1 - r.z0);
  rg.hy = (int64_t)(b.y1 - r.y0);
  rg.hx = (int64_t)(b.x1 - r.x0);
  rg.sy = r.dx;
  rg.sz = plane;
  for (uint64_t i0 = 0; i0 < nbox && rc == R3D_OK; i0 += slab) {
    const uint32_t n = (uint32_t)(nbox - i0 < slab ? nbox - i0 : slab);
    GatherArgs g{n,     (uint32_t)i0, (uint32_t)b.ncy, (uint32_t)b.ncx, b.cz0,         b.cy0,     b.cx0,
                 sz0,   sy0,          sx0,             (uint32_t)nsy,   (uint32_t)nsx, dtab,      base,
                 w.offsets, w.lengths, w.state};
    hipLaunchKernelGGL(k_shard_gather, dim3((n + 255u) / 256u), dim3(256), 0, st, g);
    if (hipGetLastError() != hipSuccess) {
      rc = r3d_fail(ctx, R3D_ERR_DEVICE, "shard_decode_region: gather launch");
      break;
    }
    r3d_blob_batch in{n, 1u, SH_CAP, 0u, base, w.offsets, w.lengths};
    r3d_blob_parse_out po{w.hdr,     w.steps,    w.pair_off, w.pair_cnt,  w.pairs,
                          w.corr_off, w.corr_cnt, w.corr_pos, w.corr_delta, w.parse_st};
    rc = r3d_hip_parse_blobs(ctx, &in, &po, stream);
    if (rc) break;
    r3d_chunk_batch cb{};
    cb.nchunks = n;
    cb.coef_format = R3D_COEF_SPARSE;
    cb.pair_offsets = w.pair_off;
    cb.pairs = w.pairs;
    cb.npairs = (uint64_t)n * SH_CAP;
    cb.headers = w.hdr;
    cb.step_tables = w.steps;
    cb.ntables = n;
    cb.corr_offsets = w.corr_off;
    cb.corr_pos = w.corr_pos;
    cb.corr_delta = w.corr_delta;
    cb.ncorr = (uint64_t)n * SH_CAP;
    cb.out = dst;
    cb.pair_counts = w.pair_cnt;
    cb.corr_counts = w.corr_cnt;
    rg.first = (uint32_t)i0;
    rc = decode_region(ctx, &cb, rg, st);
    if (rc) break;
    hipLaunchKernelGGL(k_shard_finish, dim3((n + 255u) / 256u), dim3(256), 0, st, w.state, w.parse_st, n,
                       chunk_status ? chunk_status + i0 : nullptr, counts);
    if (hipGetLastError() != hipSuccess) rc = r3d_fail(ctx, R3D_ERR_DEVICE, "shard_decode_region: finish launch");
  }
  // the staged shard table is read up to the last gather
  if (hipEventRecord(sl->ev, st) == hipSuccess) sl->recorded = true;
  return rc;
}
