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

// The entries of the filled bricks, once every part has folded its maximum.
__global__ __launch_bounds__(256) void k_bk_entries(const BkJob *jobs, uint32_t njobs, const uint32_t *maxv,
                                                    uint32_t *page) {
  const uint32_t j = blockIdx.x * 256u + threadIdx.x;
  if (j < njobs) page[jobs[j].word] = jobs[j].slot | (maxv[j] << 24);
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
  bool ok = hipMalloc(&b->d_jobs, (sizeof(BkJob) + 4u) * (uint64_t)b->max_jobs) == hipSuccess;
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
  for (uint32_t l = 0; l < R3D_MAX_LEVELS; l++)
    if (b->d_lev[l]) hipFree(b->d_lev[l]);
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
                         f.njobs, (const uint32_t *)f.maxv, f.page);
      err = hipGetLastError();
    }
  }
  const hipError_t sync = hipStreamSynchronize(st);
  R3D_HIP_CHECK(ctx, err);
  R3D_HIP_CHECK(ctx, sync);
  return R3D_OK;
}
