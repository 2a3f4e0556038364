// r3d_internal.h -- state shared by the library's translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>

#include <vector>

#include "r3d_hip.h"

#define R3D_BRICK 16u
#define R3D_BRICK_STRIDE 17u                  // 16 payload + 1 apron voxel per axis
#define R3D_BRICK_CONTENT (17u * 17u * 17u)   // 4913 bytes of voxels per brick
#define R3D_BRICK_BYTES 4928u                 // brick stride: content padded to 16 bytes
#define R3D_OCC_BLOCK 8u

namespace r3d {

// Per-level descriptor as the kernels read it (48 bytes, staged in LDS).
struct LevelDev {
  float fx, fy, fz;         // level dims as floats
  uint32_t off_lo;          // byte offset of the level's first brick: low word
  float hx, hy, hz;         // fx - 1, fy - 1, fz - 1: a footprint origin's upper clamp
  uint32_t off_hi;          // and high word
  uint32_t bx, by, bz;      // bricks along x, y and z
  uint32_t pad;
};

struct VolDev {
  const uint8_t *base;      // all levels' bricks
  const LevelDev *levels;   // device copy of the level table
  const uint8_t *occ;       // dilated occupancy [oz][oy][ox]
  uint32_t nlev;
  uint32_t maxdim;
  uint32_t ox, oy, oz;
  uint32_t narrow;          // the pyramid's bytes fit a 32-bit offset
};

struct ViewSlot {
  void *host = nullptr;   // pinned staging
  void *dev = nullptr;
  size_t cap = 0;
  hipEvent_t ev = nullptr;
  bool recorded = false;
};

}  // namespace r3d

struct r3d_hip_ctx {
  int device = 0;
  int num_cus = 0;
  char err[512] = {0};
  float basis[256];            // host copy of the DCT basis
  float *d_tf = nullptr;       // 256 x 4 FP32 transfer table
  unsigned long long *d_counters = nullptr;  // 8 x u64 scratch for stats
  uint32_t *d_partials = nullptr;            // per-tile raycast counter partials
  size_t partials_cap = 0;                   // in tiles
  uint32_t *d_order = nullptr;               // raycast launch order (bucket counters, tile codes, order)
  size_t order_cap = 0;                      // in words
  float *d_gtf = nullptr;                    // adjoint: per-block TF gradient partials
  size_t gtf_cap = 0;                        // bytes
  r3d::ViewSlot slots[4];
  int slot_next = 0;
  r3d_raycast_tuning tuning{8u, 8u, 0u, 0.0f};  // r3d_hip_raycast's launch shape
  uint32_t quality = R3D_QUALITY_FULL;           // gradient of FULL-mode views (r3d_hip_set_quality)
  uint32_t *d_queue = nullptr;                   // tile queues of r3d_hip_raycast, one per kind
  void *d_ws = nullptr;                          // workspace of r3d_hip_shard_decode_region
  size_t ws_cap = 0;
  void *d_dblim = nullptr;                       // 16 bytes: r3d_hip_deblock_u8's data-range thresholds
  void *d_sort = nullptr;                        // r3d_hip_parse_blobs' lane order and its sort
  size_t sort_cap = 0;
  // order buffers a larger batch replaced: kept until r3d_hip_destroy, since a
  // HIP graph that captured an earlier parse still reads them
  void *retired[48] = {};
  uint32_t nretired = 0;
  void *d_noalloc = nullptr;  // this device's fixed parse scratch (r3d_hip_parse_blobs_noalloc)
  // pane cache (R3D_RAYCAST_PANE_CACHE): the scene generation, bumped by every
  // call that changes what a view renders, and the key each view index was
  // last drawn under (0: draw)
  uint64_t scene_gen = 1;
  uint64_t pane_key[R3D_MAX_VIEWS_PER_CALL] = {};
  hipEvent_t ev_ray[2] = {nullptr, nullptr};  // the launches' GPU time, for stats
};

// What a view renders changed (data, transfer function, quality, launch shape).
inline void r3d_scene_changed(r3d_hip_ctx *c) { c->scene_gen++; }
// Outputs were written outside the pane cache: every cached view redraws.
inline void r3d_panes_reset(r3d_hip_ctx *c) {
  for (uint32_t i = 0; i < R3D_MAX_VIEWS_PER_CALL; i++) c->pane_key[i] = 0;
}

struct r3d_hip_volume {
  uint8_t *d_flat = nullptr;  // staging for r3d_hip_volume_decode: nchunks x 4096
  r3d_hip_ctx *ctx = nullptr;
  r3d_volume_desc desc{};
  uint32_t nlev = 0;
  uint32_t dims[R3D_MAX_LEVELS][3] = {};
  uint32_t bricks[R3D_MAX_LEVELS][3] = {};
  uint64_t level_off[R3D_MAX_LEVELS] = {};
  uint64_t bytes = 0;
  uint8_t *d_bricks = nullptr;
  r3d::LevelDev *d_levels = nullptr;
  uint32_t odim[3] = {};
  uint8_t *d_occ = nullptr;
  uint8_t *d_occ_raw = nullptr;
  bool wide = false;  // 64-bit brick offsets whatever the pyramid's size (R3D_WIDE_OFFSETS=1 at create)
};

struct r3d_hip_bricks {
  r3d_hip_ctx *ctx = nullptr;
  r3d_bricks_desc desc{};
  uint32_t dims[R3D_MAX_LEVELS][3] = {};   // level dims (x, y, z)
  uint32_t nb[R3D_MAX_LEVELS][3] = {};     // bricks per axis per level
  uint32_t off[R3D_MAX_LEVELS] = {};       // page word of each level's first entry
  uint32_t page_words = 0;
  uint32_t nslots = 0;
  uint32_t *d_page = nullptr;
  uint8_t *d_atlas = nullptr;
  // build scratch, sized at create and kept: levels 1.. as linear arrays, and
  // the job list with one max word per job (at most max_jobs)
  uint8_t *d_lev[R3D_MAX_LEVELS] = {};
  void *d_jobs = nullptr;
  uint32_t max_jobs = 0;
  // streaming (r3d_hip_bricks_stream_*): the source pyramid of the last build
  // (level 0 is the caller's volume) and the residency policy's host state
  const uint8_t *src0 = nullptr;
  bool stream_ready = false;             // bmax computed for the current build
  uint8_t *d_bmax = nullptr;             // per page entry: the brick's largest voxel
  std::vector<uint8_t> bmax;
  std::vector<uint32_t> brick_slot;      // per page entry: its slot, or ~0
  std::vector<uint32_t> brick_want;      // per page entry: the frame it was last requested in
  std::vector<uint32_t> slot_brick;      // per slot: its page entry (index after the header), or ~0
  std::vector<uint32_t> slot_use;        // per slot: the frame it was last wanted in
  struct Cand {
    float d2;
    uint32_t b, priority;
  };
  std::vector<Cand> cands;
  uint32_t frame = 0;
  bool stream_open = false;
  uint32_t lod_wanted[8] = {}, last_filled = 0, last_evicted = 0;
  uint64_t lod_requests[8] = {}, filled = 0, evicted = 0, submits = 0;
  // streaming from 3ddct shard stores (r3d_hip_bricks_build_from_shards): one
  // shard list per level (device-resident files); a brick's maximum is known
  // once a fill has decoded it
  bool from_shards = false;
  std::vector<r3d_shard_src> lev_shards[R3D_MAX_LEVELS];
  std::vector<uint8_t> bmax_known;
  uint32_t *h_maxv = nullptr;     // pinned: the maxima of the last submit's fills
  hipEvent_t ev_fill = nullptr;   // recorded after their copy
  std::vector<uint32_t> pend;     // page entries (after the header) of those fills
};

// error helpers
int r3d_fail(r3d_hip_ctx *ctx, int code, const char *fmt, ...) __attribute__((format(printf, 3, 4)));
#define R3D_HIP_CHECK(ctx, expr)                                                        \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess)                                                               \
      return r3d_fail((ctx), R3D_ERR_DEVICE, "%s: %s", #expr, hipGetErrorString(e_));   \
  } while (0)

namespace r3d {
// Workgroups are dispatched to the 8 XCDs round-robin in blockIdx order, and
// each XCD has its own L2. For a grid of n workgroups of equal cost, this map
// gives XCD k the contiguous band of work items [k n/8, (k + 1) n/8) in order
// (the remainder past 8 (n / 8) keeps its own index), so neighbouring items,
// which read each other's edge lines, run on the same XCD a few slots apart
// and find those lines in its L2.
__device__ __forceinline__ uint32_t xcd_band(uint32_t b, uint32_t n) {
  const uint32_t per = n / 8u;
  return b < 8u * per ? (b % 8u) * per + b / 8u : b;
}

// Copies the FP32 DCT basis (dct3d.c:38-49) into the decoder's constant bank.
int upload_basis(r3d_hip_ctx *ctx);
// Uploads the 3ddct scan order and the powf(1 + f, 0.65f) weights (computed by
// the host's libm) for the blob parser and encoder.
int upload_blob_tables(r3d_hip_ctx *ctx);
// Takes the device address of r3d_hip_parse_blobs_noalloc's fixed scratch
// on the context's (current) device, once per device.
int parse_noalloc_setup(r3d_hip_ctx *ctx);
// Uploads the CRC-32C byte table and the x^(2^k) mod P powers used to fold
// piecewise CRCs.
int upload_crc_tables(r3d_hip_ctx *ctx);
// Copies n bytes into the next staging slot (pinned host + device buffer);
// the caller records the slot's event after the launches that read it.
int stage(r3d_hip_ctx *ctx, const void *src, size_t n, hipStream_t s, ViewSlot **slot_out);
r3d::VolDev vol_dev(const r3d_hip_volume *v);

// A region destination for the exact decode (shard region decodes): chunk c
// of the batch is cell first + c of an ncz x ncy x ncx chunk box (x fastest) whose
// first chunk starts at (rz, ry, rx) relative to the region origin; only
// voxels in [0, hz) x [0, hy) x [0, hx) of the region are written, at row
// stride sy and slice stride sz.
struct RegionDst {
  uint32_t ncz, ncy, ncx, first;  // first: the box cell of the batch's chunk 0
  int64_t rz, ry, rx, hz, hy, hx;
  uint64_t sy, sz;
};
int decode_region(r3d_hip_ctx *ctx, const r3d_chunk_batch *b, const RegionDst &rg, hipStream_t stream);
// Grow-only device workspace of the context (shard region decodes).
int workspace(r3d_hip_ctx *ctx, size_t bytes, void **out);
// 128^3 bricks of one shard store decoded into atlas slots in one pass (one
// index check per shard, one gather per brick, one parse and one decode for
// all of them): brick (bx, by, bz) of a store of dims (nz, ny, nx), each a
// multiple of 16, into slot `slot` (z-major 128^3 u8); voxels past the store
// are zero. At most 64 bricks.
struct BrickTarget {
  uint32_t bx, by, bz, slot;
};
int shard_decode_bricks(r3d_hip_ctx *ctx, const r3d_shard_src *shards, uint32_t nshards, const uint64_t dims[3],
                        const BrickTarget *t, uint32_t n, uint8_t *atlas, hipStream_t st);
}  // namespace r3d
