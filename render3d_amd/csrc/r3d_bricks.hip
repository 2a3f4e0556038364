// r3d_bricks.hip -- bricks-LOD volumes: the page table and slot atlas of
// raycast.slang's R3D_MODE 3 with BR_LOD_FLAG (the render is in
// r3d_raycast.hip, march_bricks).
//
// The page table is the reference's layout (vkbackend.c:4185-4205), so a
// streamer written against it can fill it: a 64-word header (nlev, base
// shape, per level {entry offset, packed brick dims, scale}) and one entry
// per 128^3 brick of every level, slot | max << 24 or 0xFFFFFFFF. The atlas is
// linear HBM: slot s is 2 MiB at atlas + s * 128^3, z-major, so a trilinear
// footprint is 8 byte loads at fixed offsets.
//
// r3d_hip_bricks_build makes the pyramid with the rounded 2x box
// ((sum of 8 + 4) / 8, edge voxels repeated: tools/lodpack/main.c:216-227)
// into temporary linear levels, then one workgroup per resident brick copies
// its 128^3 voxels into its slot (zeros past the level's edge) and writes its
// entry with the brick's maximum.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <new>
#include <vector>

#include "r3d_internal.h"

namespace r3d {

constexpr uint32_t BK = 128u;
constexpr uint64_t BK_VOX = 128ull * 128ull * 128ull;

__global__ __launch_bounds__(256) void k_bk_down(const uint8_t *src, uint32_t nx, uint32_t ny, uint32_t nz,
                                                 uint8_t *dst, uint32_t mx, uint32_t my, uint32_t mz) {
  const uint64_t n = (uint64_t)mx * my * mz;
  for (uint64_t i = (uint64_t)blockIdx.x * 256u + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256u) {
    const uint32_t x = (uint32_t)(i % mx), y = (uint32_t)((i / mx) % my), z = (uint32_t)(i / ((uint64_t)mx * my));
    uint32_t sum = 4u;
#pragma unroll
    for (uint32_t d = 0; d < 8u; d++) {
      uint32_t sx = 2u * x + (d & 1u), sy = 2u * y + ((d >> 1) & 1u), sz = 2u * z + (d >> 2);
      sx = sx < nx ? sx : nx - 1u;
      sy = sy < ny ? sy : ny - 1u;
      sz = sz < nz ? sz : nz - 1u;
      sum += src[((uint64_t)sz * ny + sy) * nx + sx];
    }
    dst[i] = (uint8_t)(sum / 8u);
  }
}

// The 2x box of a level whose x extent is a multiple of 16 and whose base is
// 16-byte aligned: each lane makes 8 consecutive voxels of a destination row
// from one 16-byte load of each of its 4 source rows (rows past the source's
// y / z edge are clamped; x never needs it since 2x + 1 < nx).
__global__ __launch_bounds__(256) void k_bk_down16(const uint8_t *src, uint32_t nx, uint32_t ny, uint32_t nz,
                                                   uint8_t *dst, uint32_t mx, uint32_t my, uint32_t mz) {
  const uint32_t gx = mx / 8u;  // 8-voxel groups per destination row
  const uint64_t n = (uint64_t)gx * my * mz;
  for (uint64_t i = (uint64_t)blockIdx.x * 256u + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256u) {
    const uint32_t g = (uint32_t)(i % gx), y = (uint32_t)((i / gx) % my), z = (uint32_t)(i / ((uint64_t)gx * my));
    uint32_t acc[8];
#pragma unroll
    for (int k = 0; k < 8; k++) acc[k] = 4u;
#pragma unroll
    for (uint32_t d = 0; d < 4u; d++) {
      const uint32_t sy = min(2u * y + (d & 1u), ny - 1u), sz = min(2u * z + (d >> 1), nz - 1u);
      const uint4 q = *(const uint4 *)(src + ((uint64_t)sz * ny + sy) * nx + 16u * g);
      const uint32_t w[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
      for (int k = 0; k < 8; k++)
        acc[k] += ((w[k >> 1] >> (16 * (k & 1))) & 255u) + ((w[k >> 1] >> (16 * (k & 1) + 8)) & 255u);
    }
    uint2 o;
    o.x = (acc[0] / 8u) | (acc[1] / 8u) << 8 | (acc[2] / 8u) << 16 | (acc[3] / 8u) << 24;
    o.y = (acc[4] / 8u) | (acc[5] / 8u) << 8 | (acc[6] / 8u) << 16 | (acc[7] / 8u) << 24;
    *(uint2 *)(dst + ((uint64_t)z * my + y) * mx + 8u * g) = o;
  }
}

struct BkJob {
  uint32_t level, cx, cy, cz;
  uint32_t word, slot;
};

struct BkFill {
  const uint8_t *lev[R3D_MAX_LEVELS];
  uint32_t dims[R3D_MAX_LEVELS][3];
  const BkJob *jobs;
  uint32_t njobs;
  uint8_t *atlas;
  uint32_t *page;
  uint32_t *maxv;  // per job, zeroed before the fill
};

constexpr uint32_t BK_PARTS = 64u;  // workgroups per brick: 256 rows of 128 voxels each

// BK_PARTS workgroups per resident brick, each copying 256 of its 16,384
// rows of 128 voxels (zeros past the level's edge) and folding its maximum
// into the brick's word with an atomic max. Rows that lie whole and 16-byte
// aligned in the level go eight lanes to a row, 16 bytes each, so that a wave
// instruction touches 8 lines rather than 64; other bricks take one row per
// lane.
__global__ __launch_bounds__(256) void k_bk_fill(BkFill a) {
  __shared__ uint32_t s_max[4];
  const uint32_t j = blockIdx.x / BK_PARTS, part = blockIdx.x % BK_PARTS;
  if (j >= a.njobs) return;
  const BkJob job = a.jobs[j];
  const uint8_t *L = a.lev[job.level];
  const uint32_t nx = a.dims[job.level][0], ny = a.dims[job.level][1], nz = a.dims[job.level][2];
  uint8_t *S = a.atlas + (uint64_t)job.slot * BK_VOX;
  const uint32_t x0 = job.cx * BK;
  const bool whole_x = x0 + BK <= nx && nx % 16u == 0u && ((uintptr_t)L & 15u) == 0u;
  uint32_t mx = 0;
  if (whole_x) {
    const uint32_t c = threadIdx.x & 7u, rs = threadIdx.x >> 3;  // 16-byte column, first row
    uint4 q[8];
#pragma unroll
    for (int i = 0; i < 8; i++) {
      const uint32_t r = part * 256u + rs + 32u * (uint32_t)i, y = r & 127u, z = r >> 7;
      const uint32_t gy = job.cy * BK + y, gz = job.cz * BK + z;
      q[i] = make_uint4(0, 0, 0, 0);
      if (gy < ny && gz < nz) q[i] = *(const uint4 *)(L + ((uint64_t)gz * ny + gy) * nx + x0 + 16u * c);
    }
#pragma unroll
    for (int i = 0; i < 8; i++) {
      const uint32_t r = part * 256u + rs + 32u * (uint32_t)i, y = r & 127u, z = r >> 7;
      *(uint4 *)(S + ((uint64_t)z * BK + y) * BK + 16u * c) = q[i];
      const uint32_t w[4] = {q[i].x, q[i].y, q[i].z, q[i].w};
#pragma unroll
      for (int k = 0; k < 4; k++)
#pragma unroll
        for (int bb = 0; bb < 4; bb++) {
          const uint32_t v = (w[k] >> (8 * bb)) & 255u;
          mx = v > mx ? v : mx;
        }
    }
  } else {
    const uint32_t r = part * 256u + threadIdx.x;  // row of the brick
    const uint32_t y = r & 127u, z = r >> 7;
    const uint32_t gy = job.cy * BK + y, gz = job.cz * BK + z;
    uint8_t *d8 = S + ((uint64_t)z * BK + y) * BK;
    if (gy >= ny || gz >= nz) {
      for (int k = 0; k < 8; k++) ((uint4 *)d8)[k] = make_uint4(0, 0, 0, 0);
    } else {
      const uint8_t *row = L + ((uint64_t)gz * ny + gy) * nx + x0;
      for (uint32_t x = 0; x < BK; x++) {
        const uint32_t v = x0 + x < nx ? row[x] : 0u;
        d8[x] = (uint8_t)v;
        mx = v > mx ? v : mx;
      }
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    const uint32_t t = __shfl_xor(mx, o);
    mx = t > mx ? t : mx;
  }
  if ((threadIdx.x & 63u) == 0u) s_max[threadIdx.x >> 6] = mx;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t m = s_max[0];
    for (int w = 1; w < 4; w++) m = s_max[w] > m ? s_max[w] : m;
    atomicMax(&a.maxv[j], m);
  }
}

__global__ __launch_bounds__(256) void k_bk_zero(uint32_t *p, uint32_t n) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i < n) p[i] = 0u;
}

// The entries of the filled bricks, once every part has folded its maximum;
// a brick whose maximum is below floor stays absent (vkbackend.c:5265-5269).
__global__ __launch_bounds__(256) void k_bk_entries(const BkJob *jobs, uint32_t njobs, const uint32_t *maxv,
                                                    uint32_t *page, uint32_t floor) {
  const uint32_t j = blockIdx.x * 256u + threadIdx.x;
  if (j < njobs) page[jobs[j].word] = maxv[j] >= floor ? jobs[j].slot | (maxv[j] << 24) : 0xFFFFFFFFu;
}

// The largest voxel of each filled slot: BK_PARTS workgroups per job, 256
// rows of 128 voxels each, 16 bytes per load (a slot is 2 MiB aligned).
__global__ __launch_bounds__(256) void k_bk_slotmax(const BkJob *jobs, const uint8_t *atlas, uint32_t *maxv) {
  __shared__ uint32_t s_max[4];
  const uint32_t j = blockIdx.x / BK_PARTS, part = blockIdx.x % BK_PARTS;
  const uint4 *S = (const uint4 *)(atlas + (uint64_t)jobs[j].slot * BK_VOX + (uint64_t)part * 256u * 128u);
  uint32_t mx = 0;
  for (uint32_t i = threadIdx.x; i < 256u * 128u / 16u; i += 256u) {
    const uint4 q = S[i];
    const uint32_t w[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
    for (int k = 0; k < 16; k++) mx = max(mx, (w[k >> 2] >> (8 * (k & 3))) & 255u);
  }
  for (int o = 32; o > 0; o >>= 1) mx = max(mx, (uint32_t)__shfl_xor(mx, o));
  if ((threadIdx.x & 63u) == 0u) s_max[threadIdx.x >> 6] = mx;
  __syncthreads();
  if (threadIdx.x == 0) atomicMax(&maxv[j], max(max(s_max[0], s_max[1]), max(s_max[2], s_max[3])));
}


// The entries of evicted bricks become absent before their slots are refilled
// (queue-ordered on the stream, so a frame rendered earlier still reads the
// old entry and a later one the new).
__global__ __launch_bounds__(256) void k_bk_evict(const uint32_t *words, uint32_t n, uint32_t *page) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i < n) page[words[i]] = 0xFFFFFFFFu;
}

struct BkMax {
  const uint8_t *lev[R3D_MAX_LEVELS];
  uint32_t dims[R3D_MAX_LEVELS][3];
  uint32_t nb[R3D_MAX_LEVELS][3];
  uint32_t off[R3D_MAX_LEVELS];  // page word of each level's first entry
  uint32_t nlev;
  uint8_t *bmax;                 // per entry (word - 64)
};

// One workgroup per brick of every level: its largest voxel (the streamer's
// gate test: bricks below the gate never take a slot). Rows of 128 voxels,
// one per lane, 16 bytes at a time where the row lies whole in the level.
__global__ __launch_bounds__(256) void k_bk_bmax(BkMax a) {
  __shared__ uint32_t s_max[4];
  const uint32_t e = blockIdx.x;
  uint32_t l = 0;
  while (l + 1u < a.nlev && e + 64u >= a.off[l + 1u]) l++;
  const uint32_t k = e + 64u - a.off[l];
  const uint32_t cx = k % a.nb[l][0], cy = (k / a.nb[l][0]) % a.nb[l][1], cz = k / (a.nb[l][0] * a.nb[l][1]);
  const uint32_t nx = a.dims[l][0], ny = a.dims[l][1], nz = a.dims[l][2];
  const uint8_t *L = a.lev[l];
  const uint32_t x0 = cx * BK, x1 = min(x0 + BK, nx);
  const bool vec = x1 - x0 == BK && nx % 16u == 0u && ((uintptr_t)L & 15u) == 0u;
  uint32_t mx = 0;
  for (uint32_t r = threadIdx.x; r < BK * BK; r += 256u) {
    const uint32_t gy = cy * BK + (r & 127u), gz = cz * BK + (r >> 7);
    if (gy >= ny || gz >= nz) continue;
    const uint8_t *row = L + ((uint64_t)gz * ny + gy) * nx;
    if (vec) {
      for (uint32_t x = x0; x < x1; x += 16u) {
        const uint4 q = *(const uint4 *)(row + x);
        const uint32_t w[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
        for (int i = 0; i < 16; i++) mx = max(mx, (w[i >> 2] >> (8 * (i & 3))) & 255u);
      }
    } else {
      for (uint32_t x = x0; x < x1; x++) mx = max(mx, (uint32_t)row[x]);
    }
  }
  for (int o = 32; o > 0; o >>= 1) mx = max(mx, (uint32_t)__shfl_xor(mx, o));
  if ((threadIdx.x & 63u) == 0u) s_max[threadIdx.x >> 6] = mx;
  __syncthreads();
  if (threadIdx.x == 0) a.bmax[e] = (uint8_t)max(max(s_max[0], s_max[1]), max(s_max[2], s_max[3]));
}

}  // namespace r3d

using namespace r3d;

extern "C" int r3d_hip_bricks_create(r3d_hip_ctx *ctx, const r3d_bricks_desc *d, r3d_hip_bricks **out) {
  if (!ctx) return R3D_ERR_INVALID;
  if (!d || !out) return r3d_fail(ctx, R3D_ERR_INVALID, "bricks_create: NULL argument");
  *out = nullptr;
  if (d->nx == 0 || d->ny == 0 || d->nz == 0)
    return r3d_fail(ctx, R3D_ERR_INVALID, "bricks_create: empty shape %ux%ux%u", d->nx, d->ny, d->nz);
  if (d->nx > 1023u * BK || d->ny > 1023u * BK || d->nz > 1023u * BK)
    return r3d_fail(ctx, R3D_ERR_RANGE, "bricks_create: more than 1023 bricks per axis");
  if (d->nlev == 0 || d->nlev > 15u) return r3d_fail(ctx, R3D_ERR_INVALID, "bricks_create: nlev %u outside 1..15", d->nlev);
  if (d->abpa == 0 || d->abpa > 255u) return r3d_fail(ctx, R3D_ERR_INVALID, "bricks_create: abpa %u outside 1..255", d->abpa);
  r3d_hip_bricks *b = new (std::nothrow) r3d_hip_bricks;
  if (!b) return r3d_fail(ctx, R3D_ERR_NOMEM, "bricks_create: handle");
  b->ctx = ctx;
  b->desc = *d;
  uint64_t words = 64u;
  for (uint32_t l = 0; l < d->nlev; l++) {
    for (int a = 0; a < 3; a++) {
      const uint32_t base = a == 0 ? d->nx : a == 1 ? d->ny : d->nz;
      b->dims[l][a] = l == 0 ? base : (b->dims[l - 1][a] > 1u ? b->dims[l - 1][a] / 2u : 1u);
      b->nb[l][a] = (b->dims[l][a] + BK - 1u) / BK;
    }
    b->off[l] = (uint32_t)words;
    words += (uint64_t)b->nb[l][0] * b->nb[l][1] * b->nb[l][2];
  }
  if (words > 0xFFFFFFFFull) {
    delete b;
    return r3d_fail(ctx, R3D_ERR_RANGE, "bricks_create: page table too large");
  }
  b->page_words = (uint32_t)words;
  b->nslots = d->abpa * d->abpa * d->abpa;
  if (hipSetDevice(ctx->device) != hipSuccess || hipMalloc(&b->d_page, 4ull * b->page_words) != hipSuccess ||
      hipMalloc(&b->d_atlas, (uint64_t)b->nslots * BK_VOX) != hipSuccess) {
    r3d_hip_bricks_destroy(b);
    return r3d_fail(ctx, R3D_ERR_NOMEM, "bricks_create: %u slots of 2 MiB", d->abpa * d->abpa * d->abpa);
  }
  // the build's scratch, allocated once so that a build allocates nothing
  const uint64_t entries = words - 64u;
  b->max_jobs = (uint32_t)(entries < b->nslots ? entries : b->nslots);
  bool ok = hipMalloc(&b->d_jobs, (sizeof(BkJob) + 4u) * (uint64_t)b->max_jobs) == hipSuccess &&
            hipMalloc(&b->d_bmax, entries) == hipSuccess;
  for (uint32_t l = 1; l < d->nlev && ok; l++)
    ok = hipMalloc(&b->d_lev[l], (uint64_t)b->dims[l][0] * b->dims[l][1] * b->dims[l][2]) == hipSuccess;
  if (!ok) {
    r3d_hip_bricks_destroy(b);
    return r3d_fail(ctx, R3D_ERR_NOMEM, "bricks_create: build scratch (coarser levels and job list)");
  }
  std::vector<uint32_t> page(b->page_words, 0xFFFFFFFFu);
  for (uint32_t i = 0; i < 64u; i++) page[i] = 0u;
  page[0] = d->nlev;
  page[1] = d->nx;
  page[2] = d->ny;
  page[3] = d->nz;
  for (uint32_t l = 0; l < d->nlev; l++) {
    page[4u + l * 4u] = b->off[l];
    page[5u + l * 4u] = b->nb[l][0] | (b->nb[l][1] << 10) | (b->nb[l][2] << 20);
    page[6u + l * 4u] = 1u << l;
  }
  if (hipMemcpy(b->d_page, page.data(), 4ull * b->page_words, hipMemcpyHostToDevice) != hipSuccess) {
    r3d_hip_bricks_destroy(b);
    return r3d_fail(ctx, R3D_ERR_DEVICE, "bricks_create: page upload");
  }
  *out = b;
  return R3D_OK;
}

extern "C" void r3d_hip_bricks_destroy(r3d_hip_bricks *b) {
  if (!b) return;
  if (b->ctx) hipSetDevice(b->ctx->device);
  if (b->d_page) hipFree(b->d_page);
  if (b->d_atlas) hipFree(b->d_atlas);
  if (b->d_jobs) hipFree(b->d_jobs);
  if (b->d_bmax) hipFree(b->d_bmax);
  for (uint32_t l = 0; l < R3D_MAX_LEVELS; l++)
    if (b->d_lev[l]) hipFree(b->d_lev[l]);
  if (b->ev_fill) {
    hipEventSynchronize(b->ev_fill);
    hipEventDestroy(b->ev_fill);
  }
  if (b->h_maxv) hipHostFree(b->h_maxv);
  delete b;
}

extern "C" int r3d_hip_bricks_tables(const r3d_hip_bricks *b, uint32_t **page, uint32_t *page_words, uint8_t **atlas,
                                     uint32_t *nslots) {
  if (!b) return R3D_ERR_INVALID;
  if (page) *page = b->d_page;
  if (page_words) *page_words = b->page_words;
  if (atlas) *atlas = b->d_atlas;
  if (nslots) *nslots = b->nslots;
  return R3D_OK;
}

extern "C" int r3d_hip_bricks_build(r3d_hip_ctx *ctx, r3d_hip_bricks *b, const uint8_t *vol, const uint8_t *resident,
                                    void *stream) {
  if (!ctx) return R3D_ERR_INVALID;
  if (!b || b->ctx != ctx || !vol) return r3d_fail(ctx, R3D_ERR_INVALID, "bricks_build: bad handle or volume");
  R3D_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  hipStream_t st = (hipStream_t)stream;
  const uint32_t nlev = b->desc.nlev;
  // resident bricks in page order, slots in that order
  std::vector<BkJob> jobs;
  std::vector<uint32_t> page(b->page_words - 64u, 0xFFFFFFFFu);
  for (uint32_t l = 0; l < nlev; l++) {
    const uint32_t n = b->nb[l][0] * b->nb[l][1] * b->nb[l][2];
    for (uint32_t k = 0; k < n; k++) {
      const uint32_t w = b->off[l] + k;
      if (resident && !resident[w - 64u]) continue;
      if (jobs.size() >= b->nslots)
        return r3d_fail(ctx, R3D_ERR_RANGE, "bricks_build: more resident bricks than the %u slots", b->nslots);
      jobs.push_back(BkJob{l, k % b->nb[l][0], (k / b->nb[l][0]) % b->nb[l][1], k / (b->nb[l][0] * b->nb[l][1]), w,
                           (uint32_t)jobs.size()});
    }
  }
  // levels 1.. as linear arrays in the handle's scratch
  BkFill f{};
  f.lev[0] = vol;
  for (uint32_t l = 0; l < nlev; l++)
    for (int a = 0; a < 3; a++) f.dims[l][a] = b->dims[l][a];
  for (uint32_t l = 1; l < nlev; l++) {
    uint8_t *p = b->d_lev[l];
    f.lev[l] = p;
    const uint64_t n = (uint64_t)b->dims[l][0] * b->dims[l][1] * b->dims[l][2];
    const uint32_t snx = b->dims[l - 1][0], mxl = b->dims[l][0];
    const bool vec16 = snx % 16u == 0u && mxl * 2u == snx && ((uintptr_t)f.lev[l - 1] & 15u) == 0u;
    const uint64_t items = vec16 ? n / 8u : n;
    const uint64_t blocks = (items + 255u) / 256u;
    const uint32_t grid = (uint32_t)(blocks < 65536u ? blocks : 65536u);
    if (vec16)
      hipLaunchKernelGGL(k_bk_down16, dim3(grid), dim3(256), 0, st, f.lev[l - 1], snx, b->dims[l - 1][1],
                         b->dims[l - 1][2], p, mxl, b->dims[l][1], b->dims[l][2]);
    else
      hipLaunchKernelGGL(k_bk_down, dim3(grid), dim3(256), 0, st, f.lev[l - 1], snx, b->dims[l - 1][1],
                         b->dims[l - 1][2], p, mxl, b->dims[l][1], b->dims[l][2]);
    R3D_HIP_CHECK(ctx, hipGetLastError());
  }
  // the entries reset to absent, then the resident ones written by the fill;
  // the copies are ordered on the stream after work that may read the page.
  // Once a copy is queued, every exit synchronizes first: the copies read the
  // host's page and job vectors, which end with this call.
  R3D_HIP_CHECK(ctx, hipMemcpyAsync(b->d_page + 64, page.data(), 4ull * page.size(), hipMemcpyHostToDevice, st));
  hipError_t err = hipSuccess;
  if (!jobs.empty()) {
    BkJob *djobs = (BkJob *)b->d_jobs;
    err = hipMemcpyAsync(djobs, jobs.data(), sizeof(BkJob) * jobs.size(), hipMemcpyHostToDevice, st);
    if (err == hipSuccess) {
      f.jobs = djobs;
      f.njobs = (uint32_t)jobs.size();
      f.atlas = b->d_atlas;
      f.page = b->d_page;
      f.maxv = (uint32_t *)(djobs + b->max_jobs);
      hipLaunchKernelGGL(k_bk_zero, dim3((f.njobs + 255u) / 256u), dim3(256), 0, st, f.maxv, f.njobs);
      hipLaunchKernelGGL(k_bk_fill, dim3(f.njobs * BK_PARTS), dim3(256), 0, st, f);
      hipLaunchKernelGGL(k_bk_entries, dim3((f.njobs + 255u) / 256u), dim3(256), 0, st, (const BkJob *)djobs,
                         f.njobs, (const uint32_t *)f.maxv, f.page, 0u);
      err = hipGetLastError();
    }
  }
  const hipError_t sync = hipStreamSynchronize(st);
  r3d_scene_changed(ctx);
  R3D_HIP_CHECK(ctx, err);
  R3D_HIP_CHECK(ctx, sync);
  // the streamer's view of the table: the built residents in their slots
  const uint32_t nent = b->page_words - 64u;
  b->src0 = vol;
  b->from_shards = false;
  b->pend.clear();
  b->stream_ready = false;
  b->stream_open = false;
  b->frame = 0;
  b->brick_slot.assign(nent, 0xFFFFFFFFu);
  b->brick_want.assign(nent, 0u);
  b->slot_brick.assign(b->nslots, 0xFFFFFFFFu);
  b->slot_use.assign(b->nslots, 0u);
  for (const BkJob &j : jobs) {
    b->brick_slot[j.word - 64u] = j.slot;
    b->slot_brick[j.slot] = j.word - 64u;
  }
  return R3D_OK;
}

// ---- streaming residency ----------------------------------------------------
// The bricks pump of src/vk/vkbackend.c:5244-5599 (r3d_bricks_stream_begin,
// _box, _point, _submit and the cone collect r3d_bricks_stream) over this
// handle's page table. The warm tier is the pyramid of the last build, kept
// in HBM (level 0 is the caller's volume, levels 1.. the handle's scratch),
// so a requested brick is always available and is filled on the GPU at
// submit, queue-ordered on the stream after the evictions it causes: no
// worker thread and no drain. The policy is the reference's: per frame,
// candidates are the requested bricks that are not resident (parent before
// child, priority 0 before 1, then nearest first), bricks whose maximum is
// below the gate (at least 5, the codec noise floor) never take a slot, the
// coarsest level's resident bricks are kept, a free slot is taken first, else
// the least recently wanted slot not wanted this frame is evicted, and a
// submit fills at most min(budget, 32) bricks.

namespace {

constexpr uint32_t BR_INV = 0xFFFFFFFFu;
constexpr uint32_t BR_MAX_BATCH = 32u;  // vkbackend.c:2349
constexpr int BR_NOISE_FLOOR = 5;       // vkbackend.c:2348

uint32_t base_maxdim(const r3d_hip_bricks *b) {
  return std::max(b->desc.nx, std::max(b->desc.ny, b->desc.nz));
}

// Bricks below it never take a slot: max(gate * 255 rounded, the noise floor).
int gate8(float gate) {
  const int g8 = (int)(gate * 255.0f + 0.5f);
  return g8 < BR_NOISE_FLOOR ? BR_NOISE_FLOOR : g8;
}

// r3d_lod_pick (src/core/lod.h:8-17).
uint32_t lod_pick(float distance, float pixel_cone, float base_voxels, uint32_t nlevels) {
  if (!nlevels) return 0;
  const float footprint = fmaxf(distance * pixel_cone * base_voxels, 1.0f);
  int level = (int)floorf(log2f(footprint));
  if (level < 0) level = 0;
  if (level >= (int)nlevels) level = (int)nlevels - 1;
  return (uint32_t)level;
}

// Bricks along one axis whose centres lie within radius of eye.
void axis_bounds(float eye, float radius, float edge, uint32_t n, uint32_t *lo, uint32_t *hi) {
  int a = (int)ceilf((eye - radius) / edge - 0.5f);
  int c = (int)floorf((eye + radius) / edge - 0.5f) + 1;
  a = a < 0 ? 0 : (a > (int)n ? (int)n : a);
  c = c < 0 ? 0 : (c > (int)n ? (int)n : c);
  *lo = (uint32_t)a;
  *hi = (uint32_t)c;
}

uint32_t entry_of(const r3d_hip_bricks *b, uint32_t l, uint32_t bx, uint32_t by, uint32_t bz) {
  return b->off[l] - 64u + (bz * b->nb[l][1] + by) * b->nb[l][0] + bx;
}

// Keeps the coarsest level's resident bricks from eviction (their LRU stamps
// follow the frame).
void touch_coarsest(r3d_hip_bricks *b) {
  const uint32_t l = b->desc.nlev - 1u, n = b->nb[l][0] * b->nb[l][1] * b->nb[l][2];
  for (uint32_t i = 0; i < n; i++) {
    const uint32_t s = b->brick_slot[b->off[l] - 64u + i];
    if (s != BR_INV) b->slot_use[s] = b->frame;
  }
}

// A requested brick: resident ones are stamped as wanted, the others become
// candidates once per frame.
void candidate(r3d_hip_bricks *b, uint32_t e, float d2, uint32_t priority, int g8) {
  if (b->bmax_known[e] && (int)b->bmax[e] < g8) return;  // an unknown maximum is not gated
  const uint32_t s = b->brick_slot[e];
  if (s != BR_INV) {
    b->slot_use[s] = b->frame;
    return;
  }
  if (b->brick_want[e] == b->frame) return;
  b->brick_want[e] = b->frame;
  b->cands.push_back({d2, e, priority});
}

// A free slot, else the least recently wanted one not wanted this frame.
uint32_t pick_slot(const r3d_hip_bricks *b) {
  uint32_t best = BR_INV, bu = 0xFFFFFFFFu;
  for (uint32_t s = 0; s < b->nslots; s++) {
    if (b->slot_brick[s] == BR_INV) return s;
    if (b->slot_use[s] != b->frame && b->slot_use[s] < bu) {
      bu = b->slot_use[s];
      best = s;
    }
  }
  return best;
}

void count_wanted(r3d_hip_bricks *b, uint32_t level) {
  if (level < 8u) {
    b->lod_wanted[level]++;
    b->lod_requests[level]++;
  }
}

// The per-brick maxima of the built pyramid (once per build).
int stream_ready(r3d_hip_bricks *b) {
  if (b->stream_ready || b->from_shards) return R3D_OK;
  r3d_hip_ctx *ctx = b->ctx;
  if (!b->src0) return r3d_fail(ctx, R3D_ERR_INVALID, "bricks_stream: build the handle first");
  R3D_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  BkMax a{};
  a.lev[0] = b->src0;
  a.nlev = b->desc.nlev;
  for (uint32_t l = 0; l < a.nlev; l++) {
    if (l) a.lev[l] = b->d_lev[l];
    for (int k = 0; k < 3; k++) {
      a.dims[l][k] = b->dims[l][k];
      a.nb[l][k] = b->nb[l][k];
    }
    a.off[l] = b->off[l];
  }
  a.bmax = b->d_bmax;
  const uint32_t nent = b->page_words - 64u;
  hipLaunchKernelGGL(k_bk_bmax, dim3(nent), dim3(256), 0, 0, a);
  R3D_HIP_CHECK(ctx, hipGetLastError());
  b->bmax.assign(nent, 0);
  R3D_HIP_CHECK(ctx, hipMemcpy(b->bmax.data(), b->d_bmax, nent, hipMemcpyDeviceToHost));
  b->bmax_known.assign(nent, 1u);
  b->stream_ready = true;
  return R3D_OK;
}

// Decodes each job's brick from its level's shard store straight into its
// slot: the brick's 128^3 box of its level, zeros past the level's edge and
// in missing or corrupt chunks (the reference reader, shardio.c:269-306).
// The bricks of a level go in one batched pass (one index check per shard,
// one parse and one decode for all of them); a level whose dims are not
// multiples of 16 takes one region decode per brick.
int shard_fill(r3d_hip_ctx *ctx, r3d_hip_bricks *b, const BkJob *jobs, uint32_t n, void *stream) {
  for (uint32_t l = 0; l < b->desc.nlev; l++) {
    const std::vector<r3d_shard_src> &sh = b->lev_shards[l];
    const uint64_t dims[3] = {b->dims[l][2], b->dims[l][1], b->dims[l][0]};
    const bool batched = dims[0] % 16u == 0u && dims[1] % 16u == 0u && dims[2] % 16u == 0u;
    std::vector<BrickTarget> tg;
    for (uint32_t j = 0; j < n; j++) {
      const BkJob &q = jobs[j];
      if (q.level != l) continue;
      if (batched) {
        tg.push_back(BrickTarget{q.cx, q.cy, q.cz, q.slot});
        continue;
      }
      r3d_shard_region r{};
      r.nz = dims[0];
      r.ny = dims[1];
      r.nx = dims[2];
      r.z0 = (uint64_t)q.cz * BK;
      r.y0 = (uint64_t)q.cy * BK;
      r.x0 = (uint64_t)q.cx * BK;
      r.dz = r.dy = r.dx = BK;
      const int rc = r3d_hip_shard_decode_region(ctx, sh.empty() ? nullptr : sh.data(), (uint32_t)sh.size(), &r,
                                                 b->d_atlas + (uint64_t)q.slot * BK_VOX, nullptr, nullptr, stream);
      if (rc != R3D_OK) return rc;
    }
    for (size_t i = 0; i < tg.size(); i += 64u) {
      const uint32_t m = (uint32_t)(tg.size() - i < 64u ? tg.size() - i : 64u);
      const int rc = shard_decode_bricks(ctx, sh.empty() ? nullptr : sh.data(), (uint32_t)sh.size(), dims,
                                         tg.data() + i, m, b->d_atlas, (hipStream_t)stream);
      if (rc != R3D_OK) return rc;
    }
  }
  return R3D_OK;
}

}  // namespace

// The handle's source as one 3ddct shard store per level (the level's dims are
// the store's), for streaming without a decoded pyramid: the bricks whose
// resident flag is set are decoded into slots in page order, the others
// start absent with unknown maxima.
extern "C" int r3d_hip_bricks_build_from_shards(r3d_hip_ctx *ctx, r3d_hip_bricks *b,
                                                const r3d_bricks_level_src *levels, uint32_t nlev,
                                                const uint8_t *resident, void *stream) {
  if (!ctx) return R3D_ERR_INVALID;
  if (!b || b->ctx != ctx || !levels)
    return r3d_fail(ctx, R3D_ERR_INVALID, "bricks_build_from_shards: bad handle or level list");
  if (nlev != b->desc.nlev)
    return r3d_fail(ctx, R3D_ERR_INVALID, "bricks_build_from_shards: %u level stores for %u levels", nlev,
                    b->desc.nlev);
  for (uint32_t l = 0; l < nlev; l++) {
    if (levels[l].nshards && !levels[l].shards)
      return r3d_fail(ctx, R3D_ERR_INVALID, "bricks_build_from_shards: level %u has no shard list", l);
    for (uint32_t k = 0; k < levels[l].nshards; k++)
      if (levels[l].shards[k].nbytes && !levels[l].shards[k].bytes)
        return r3d_fail(ctx, R3D_ERR_INVALID, "bricks_build_from_shards: level %u shard %u has no bytes", l, k);
  }
  R3D_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  hipStream_t st = (hipStream_t)stream;
  const uint32_t nent = b->page_words - 64u;
  std::vector<BkJob> jobs;
  for (uint32_t l = 0; l < nlev; l++) {
    const uint32_t n = b->nb[l][0] * b->nb[l][1] * b->nb[l][2];
    for (uint32_t k = 0; k < n; k++) {
      const uint32_t w = b->off[l] + k;
      if (resident && !resident[w - 64u]) continue;
      if (jobs.size() >= b->nslots)
        return r3d_fail(ctx, R3D_ERR_RANGE, "bricks_build_from_shards: more resident bricks than the %u slots",
                        b->nslots);
      jobs.push_back(BkJob{l, k % b->nb[l][0], (k / b->nb[l][0]) % b->nb[l][1], k / (b->nb[l][0] * b->nb[l][1]), w,
                           (uint32_t)jobs.size()});
    }
  }
  if (!b->h_maxv && hipHostMalloc((void **)&b->h_maxv, 4ull * (b->max_jobs > 32u ? b->max_jobs : 32u)) != hipSuccess) {
    b->h_maxv = nullptr;
    return r3d_fail(ctx, R3D_ERR_NOMEM, "bricks_build_from_shards: pinned maxima");
  }
  if (!b->ev_fill && hipEventCreate(&b->ev_fill) != hipSuccess) {
    b->ev_fill = nullptr;
    return r3d_fail(ctx, R3D_ERR_DEVICE, "bricks_build_from_shards: event");
  }
  // wait for an earlier submit's copy into h_maxv before reusing it
  R3D_HIP_CHECK(ctx, hipEventSynchronize(b->ev_fill));
  for (uint32_t l = 0; l < nlev; l++)
    b->lev_shards[l].assign(levels[l].shards, levels[l].shards + levels[l].nshards);
  for (uint32_t l = nlev; l < R3D_MAX_LEVELS; l++) b->lev_shards[l].clear();
  b->from_shards = true;
  b->src0 = nullptr;
  b->stream_ready = false;
  b->stream_open = false;
  b->pend.clear();
  b->frame = 0;
  b->bmax.assign(nent, 0);
  b->bmax_known.assign(nent, 0);
  b->brick_slot.assign(nent, 0xFFFFFFFFu);
  b->brick_want.assign(nent, 0u);
  b->slot_brick.assign(b->nslots, 0xFFFFFFFFu);
  b->slot_use.assign(b->nslots, 0u);
  std::vector<uint32_t> page(nent, 0xFFFFFFFFu);
  R3D_HIP_CHECK(ctx, hipMemcpyAsync(b->d_page + 64, page.data(), 4ull * nent, hipMemcpyHostToDevice, st));
  int rc = R3D_OK;
  const uint32_t n = (uint32_t)jobs.size();
  if (n) {
    rc = shard_fill(ctx, b, jobs.data(), n, stream);
    BkJob *djobs = (BkJob *)b->d_jobs;
    uint32_t *maxv = (uint32_t *)(djobs + b->max_jobs);
    if (rc == R3D_OK && hipMemcpyAsync(djobs, jobs.data(), sizeof(BkJob) * n, hipMemcpyHostToDevice, st) != hipSuccess)
      rc = r3d_fail(ctx, R3D_ERR_DEVICE, "bricks_build_from_shards: job upload");
    if (rc == R3D_OK) {
      hipLaunchKernelGGL(k_bk_zero, dim3((n + 255u) / 256u), dim3(256), 0, st, maxv, n);
      hipLaunchKernelGGL(k_bk_slotmax, dim3(n * BK_PARTS), dim3(256), 0, st, (const BkJob *)djobs,
                         (const uint8_t *)b->d_atlas, maxv);
      hipLaunchKernelGGL(k_bk_entries, dim3((n + 255u) / 256u), dim3(256), 0, st, (const BkJob *)djobs, n,
                         (const uint32_t *)maxv, b->d_page, (uint32_t)BR_NOISE_FLOOR);
      if (hipGetLastError() != hipSuccess) rc = r3d_fail(ctx, R3D_ERR_DEVICE, "bricks_build_from_shards: launch");
    }
    if (rc == R3D_OK && hipMemcpyAsync(b->h_maxv, maxv, 4ull * n, hipMemcpyDeviceToHost, st) != hipSuccess)
      rc = r3d_fail(ctx, R3D_ERR_DEVICE, "bricks_build_from_shards: maxima copy");
  }
  // the page and job copies read host vectors that end with this call
  const hipError_t sync = hipStreamSynchronize(st);
  r3d_scene_changed(ctx);
  if (rc != R3D_OK) return rc;
  R3D_HIP_CHECK(ctx, sync);
  for (const BkJob &j : jobs) {
    const uint32_t e = j.word - 64u, idx = (uint32_t)(&j - jobs.data());
    b->brick_slot[e] = j.slot;
    b->slot_brick[j.slot] = e;
    b->bmax[e] = (uint8_t)b->h_maxv[idx];
    b->bmax_known[e] = 1u;
  }
  return R3D_OK;
}

extern "C" int r3d_hip_bricks_stream_begin(r3d_hip_bricks *b) {
  if (!b || !b->ctx) return 0;
  b->stream_open = false;
  b->last_filled = b->last_evicted = 0;
  b->cands.clear();
  if (stream_ready(b) != R3D_OK) return 0;
  if (!b->pend.empty()) {  // the last fills' maxima (vkbackend.c:5256-5282): applied once they are back
    const hipError_t q = hipEventQuery(b->ev_fill);
    if (q == hipErrorNotReady) {
      b->last_filled = (uint32_t)b->pend.size();  // still in flight: no collect this frame
      return 0;
    }
    if (q != hipSuccess) return 0;
    for (size_t j = 0; j < b->pend.size(); j++) {
      b->bmax[b->pend[j]] = (uint8_t)b->h_maxv[j];
      b->bmax_known[b->pend[j]] = 1u;
    }
    b->pend.clear();
  }
  b->frame++;
  memset(b->lod_wanted, 0, sizeof b->lod_wanted);
  b->stream_open = true;
  return 1;
}

extern "C" void r3d_hip_bricks_stream_box(r3d_hip_bricks *b, const float lo[3], const float hi[3], float pixel_cone,
                                          float gate) {
  if (!b || !b->stream_open || !lo || !hi) return;
  const int g8 = gate8(gate);
  const uint32_t nlev = b->desc.nlev;
  const float maxdim = (float)base_maxdim(b);
  const float vpp = fmaxf(pixel_cone * maxdim, 1.0f);  // base voxels per pixel
  uint32_t desired = (uint32_t)floorf(log2f(vpp));
  if (desired >= nlev) desired = nlev - 1u;
  touch_coarsest(b);
  uint32_t rng[3][2];
  float edge;
  for (;;) {  // coarser until the box holds at most 2,048 bricks
    edge = (float)(BK << desired) / maxdim;
    uint64_t count = 1;
    for (int a = 0; a < 3; a++) {
      int64_t a0 = (int64_t)floorf(lo[a] / edge), a1 = (int64_t)floorf(hi[a] / edge) + 1;
      a0 = a0 < 0 ? 0 : a0;
      a1 = a1 > (int64_t)b->nb[desired][a] ? (int64_t)b->nb[desired][a] : a1;
      a1 = a1 < a0 ? a0 : a1;
      rng[a][0] = (uint32_t)a0;
      rng[a][1] = (uint32_t)a1;
      count *= (uint64_t)(a1 - a0);
    }
    if (count <= 2048u || desired + 1u >= nlev) break;
    desired++;
  }
  const float cx0 = (lo[0] + hi[0]) * 0.5f, cy0 = (lo[1] + hi[1]) * 0.5f, cz0 = (lo[2] + hi[2]) * 0.5f;
  const bool parent = desired + 1u < nlev;
  for (uint32_t bz = rng[2][0]; bz < rng[2][1]; bz++)
    for (uint32_t by = rng[1][0]; by < rng[1][1]; by++)
      for (uint32_t bx = rng[0][0]; bx < rng[0][1]; bx++) {
        const float dx = ((float)bx + 0.5f) * edge - cx0;
        const float dy = ((float)by + 0.5f) * edge - cy0;
        const float dz = ((float)bz + 0.5f) * edge - cz0;
        const float d2 = dx * dx + dy * dy + dz * dz;
        count_wanted(b, desired);
        if (parent) candidate(b, entry_of(b, desired + 1u, bx >> 1u, by >> 1u, bz >> 1u), d2, 0u, g8);
        candidate(b, entry_of(b, desired, bx, by, bz), d2, 1u, g8);
      }
}

extern "C" void r3d_hip_bricks_stream_point(r3d_hip_bricks *b, const float p[3], uint32_t level, float gate) {
  if (!b || !b->stream_open || !p) return;
  const int g8 = gate8(gate);
  const uint32_t nlev = b->desc.nlev;
  if (level >= nlev) level = nlev - 1u;
  const float edge = (float)(BK << level) / (float)base_maxdim(b);
  uint32_t bc[3];
  for (int a = 0; a < 3; a++) {
    int64_t i = (int64_t)(p[a] / edge);
    i = i < 0 ? 0 : (i >= (int64_t)b->nb[level][a] ? (int64_t)b->nb[level][a] - 1 : i);
    bc[a] = (uint32_t)i;
  }
  if (level + 1u < nlev) candidate(b, entry_of(b, level + 1u, bc[0] >> 1u, bc[1] >> 1u, bc[2] >> 1u), 0.0f, 0u, g8);
  count_wanted(b, level);
  candidate(b, entry_of(b, level, bc[0], bc[1], bc[2]), 0.0f, 1u, g8);
}

extern "C" int r3d_hip_bricks_stream_submit(r3d_hip_ctx *ctx, r3d_hip_bricks *b, uint32_t budget, void *stream) {
  if (!ctx) return R3D_ERR_INVALID;
  if (!b || b->ctx != ctx) return r3d_fail(ctx, R3D_ERR_INVALID, "bricks_stream_submit: bad handle");
  if (!b->stream_open) return R3D_OK;
  b->stream_open = false;
  if (b->cands.empty() || budget == 0) {
    b->cands.clear();
    return R3D_OK;
  }
  if (budget > BR_MAX_BATCH) budget = BR_MAX_BATCH;
  // parent (priority 0) before child, then nearest first; equal keys keep
  // their collect order
  std::stable_sort(b->cands.begin(), b->cands.end(), [](const r3d_hip_bricks::Cand &x, const r3d_hip_bricks::Cand &y) {
    return x.priority != y.priority ? x.priority < y.priority : x.d2 < y.d2;
  });
  // staged for the device: [evicted words][jobs]
  uint32_t evict[BR_MAX_BATCH];
  BkJob jobs[BR_MAX_BATCH];
  uint32_t n = 0, nevict = 0;
  for (size_t k = 0; k < b->cands.size() && n < budget; k++) {
    const uint32_t e = b->cands[k].b;
    if (b->brick_slot[e] != BR_INV) continue;  // requested by two collects
    const uint32_t s = pick_slot(b);
    if (s == BR_INV) break;  // every slot is wanted this frame
    const uint32_t old = b->slot_brick[s];
    if (old != BR_INV) {
      b->brick_slot[old] = BR_INV;
      evict[nevict++] = old + 64u;
    }
    b->slot_brick[s] = e;
    b->slot_use[s] = b->frame;
    b->brick_slot[e] = s;
    uint32_t l = 0;
    while (l + 1u < b->desc.nlev && e + 64u >= b->off[l + 1u]) l++;
    const uint32_t k2 = e + 64u - b->off[l];
    jobs[n++] = BkJob{l, k2 % b->nb[l][0], (k2 / b->nb[l][0]) % b->nb[l][1], k2 / (b->nb[l][0] * b->nb[l][1]),
                      e + 64u, s};
  }
  b->cands.clear();
  if (!n) return R3D_OK;
  r3d_scene_changed(ctx);  // residency changes what bricks-LOD views render
  R3D_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  hipStream_t st = (hipStream_t)stream;
  // shard stores: decode into the slots first (the region decodes stage their
  // own tables), then the jobs go up; all of it is ordered on the stream, so
  // no render on it sees a slot between its eviction and its entry
  if (b->from_shards) {
    const int rc = shard_fill(ctx, b, jobs, n, stream);
    if (rc != R3D_OK) return rc;
  }
  uint8_t buf[sizeof evict + sizeof jobs];
  memcpy(buf, evict, sizeof evict);
  memcpy(buf + sizeof evict, jobs, sizeof(BkJob) * n);
  ViewSlot *slot = nullptr;
  int rc = stage(ctx, buf, sizeof evict + sizeof(BkJob) * n, st, &slot);
  if (rc != R3D_OK) return rc;
  const uint32_t *d_evict = (const uint32_t *)slot->dev;
  const BkJob *d_jobs = (const BkJob *)((const uint8_t *)slot->dev + sizeof evict);
  if (nevict) hipLaunchKernelGGL(k_bk_evict, dim3(1), dim3(256), 0, st, d_evict, nevict, b->d_page);
  uint32_t *maxv = (uint32_t *)((BkJob *)b->d_jobs + b->max_jobs);
  hipLaunchKernelGGL(k_bk_zero, dim3(1), dim3(256), 0, st, maxv, n);
  if (b->from_shards) {
    // the maxima: the gate learns them (applied at the next begin), and a
    // brick below the noise floor is not published (vkbackend.c:5265-5269)
    hipLaunchKernelGGL(k_bk_slotmax, dim3(n * BK_PARTS), dim3(256), 0, st, d_jobs, (const uint8_t *)b->d_atlas, maxv);
    hipLaunchKernelGGL(k_bk_entries, dim3(1), dim3(256), 0, st, d_jobs, n, (const uint32_t *)maxv, b->d_page,
                       (uint32_t)BR_NOISE_FLOOR);
    R3D_HIP_CHECK(ctx, hipGetLastError());
    R3D_HIP_CHECK(ctx, hipMemcpyAsync(b->h_maxv, maxv, 4ull * n, hipMemcpyDeviceToHost, st));
    R3D_HIP_CHECK(ctx, hipEventRecord(b->ev_fill, st));
    b->pend.resize(n);
    for (uint32_t j = 0; j < n; j++) b->pend[j] = jobs[j].word - 64u;
  } else {
    BkFill f{};
    f.lev[0] = b->src0;
    for (uint32_t l = 0; l < b->desc.nlev; l++) {
      if (l) f.lev[l] = b->d_lev[l];
      for (int a = 0; a < 3; a++) f.dims[l][a] = b->dims[l][a];
    }
    f.jobs = d_jobs;
    f.njobs = n;
    f.atlas = b->d_atlas;
    f.page = b->d_page;
    f.maxv = maxv;
    hipLaunchKernelGGL(k_bk_fill, dim3(n * BK_PARTS), dim3(256), 0, st, f);
    hipLaunchKernelGGL(k_bk_entries, dim3(1), dim3(256), 0, st, d_jobs, n, (const uint32_t *)f.maxv, f.page, 0u);
    R3D_HIP_CHECK(ctx, hipGetLastError());
  }
  R3D_HIP_CHECK(ctx, hipEventRecord(slot->ev, st));
  slot->recorded = true;
  b->last_filled = n;
  b->last_evicted = nevict;
  b->filled += n;
  b->evicted += nevict;
  b->submits++;
  return R3D_OK;
}

extern "C" int r3d_hip_bricks_stream(r3d_hip_ctx *ctx, r3d_hip_bricks *b, const float eye[3], const float fwd[3],
                                     float half_tan, float pixel_cone, uint32_t slice_z0, uint32_t slice_depth,
                                     float gate, uint32_t budget, void *stream) {
  if (!ctx) return R3D_ERR_INVALID;
  if (!b || b->ctx != ctx || !eye || !fwd) return r3d_fail(ctx, R3D_ERR_INVALID, "bricks_stream: bad argument");
  if (!r3d_hip_bricks_stream_begin(b)) return b->src0 ? R3D_ERR_DEVICE : r3d_fail(ctx, R3D_ERR_INVALID,
                                                                                  "bricks_stream: build first");
  if (budget == 0) {
    b->stream_open = false;
    return R3D_OK;
  }
  const float tanw = half_tan * 1.15f + 1e-3f;  // the cone, slightly widened
  const int g8 = gate8(gate);
  const uint32_t nlev = b->desc.nlev;
  touch_coarsest(b);
  const float maxdim = (float)base_maxdim(b);
  const float lod_factor = fmaxf(pixel_cone * maxdim, 1e-6f);
  // level li is wanted only inside its distance shell: walk that shell's box
  // (the coarsest level is resident and kept)
  for (uint32_t li = 0; li + 1u < nlev; li++) {
    const float edge = (float)(BK << li) / maxdim;
    const float brad = 0.8660254f * edge;
    const float outer = exp2f((float)li + 1.0f) / lod_factor + brad;
    uint32_t x0, x1, y0, y1, z0, z1;
    axis_bounds(eye[0], outer, edge, b->nb[li][0], &x0, &x1);
    axis_bounds(eye[1], outer, edge, b->nb[li][1], &y0, &y1);
    axis_bounds(eye[2], outer, edge, b->nb[li][2], &z0, &z1);
    if (slice_depth) {
      uint64_t zlast = (uint64_t)slice_z0 + slice_depth;
      if (zlast > b->desc.nz) zlast = b->desc.nz;
      const uint32_t zb = BK << li;
      const uint32_t sz0 = slice_z0 / zb, sz1 = (uint32_t)((zlast + zb - 1u) / zb);
      z0 = z0 < sz0 ? sz0 : z0;
      z1 = z1 > sz1 ? sz1 : z1;
    }
    for (uint32_t bz = z0; bz < z1; bz++)
      for (uint32_t by = y0; by < y1; by++)
        for (uint32_t bx = x0; bx < x1; bx++) {
          const float cx = ((float)bx + 0.5f) * edge - eye[0];
          const float cy = ((float)by + 0.5f) * edge - eye[1];
          const float cz = ((float)bz + 0.5f) * edge - eye[2];
          const float d2 = cx * cx + cy * cy + cz * cz;
          const float along = cx * fwd[0] + cy * fwd[1] + cz * fwd[2];
          bool vis = d2 < (edge + brad) * (edge + brad);
          if (!vis && along > 0.0f) {
            const float perp2 = d2 - along * along;
            const float rad = along * tanw + brad;
            vis = perp2 < rad * rad;
          }
          if (!vis) continue;
          const uint32_t desired = lod_pick(sqrtf(d2), pixel_cone, maxdim, nlev);
          if (li != desired) continue;
          count_wanted(b, desired);
          // the parent first: a stable, spatially matching fallback while refining
          candidate(b, entry_of(b, li + 1u, bx >> 1u, by >> 1u, bz >> 1u), d2, 0u, g8);
          candidate(b, entry_of(b, li, bx, by, bz), d2, 1u, g8);
        }
  }
  return r3d_hip_bricks_stream_submit(ctx, b, budget, stream);
}

extern "C" int r3d_hip_bricks_stream_stats(const r3d_hip_bricks *b, r3d_bricks_stream_stats *st) {
  if (!b || !st) return R3D_ERR_INVALID;
  memset(st, 0, sizeof *st);
  st->nb = b->page_words - 64u;
  st->hot_cap = b->nslots;
  for (uint32_t s = 0; s < b->nslots && s < b->slot_brick.size(); s++) st->hot += b->slot_brick[s] != BR_INV;
  st->inflight = b->last_filled;
  st->evicted_last = b->last_evicted;
  st->decoded = b->filled;
  st->evicted = b->evicted;
  st->jobs = b->submits;
  st->nlevels = b->desc.nlev;
  for (int l = 0; l < 8; l++) {
    st->lod_wanted[l] = b->lod_wanted[l];
    st->lod_requests[l] = b->lod_requests[l];
  }
  return R3D_OK;
}
