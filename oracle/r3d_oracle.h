/* r3d_oracle.h -- TEST INFRASTRUCTURE ONLY: the CPU oracle.
 *
 * A plain-C restatement of the reference's two hot paths, used as the checker
 * for the HIP path and as the timed CPU baseline. Only tests/, the smoke check
 * and bench.py's cpu_baseline leg may load liboracle.so; the product never
 * links or calls it.
 *
 *  - 3ddct decode back half (tools/3ddct/dct3d.c:38-154, 542-593, 625-639):
 *    dequantise, separable inverse DCT-III, add dc, sparse corrections,
 *    denormalise, roundf + clamp to u8. Pinned FP32 sequence: no fused
 *    multiply-add anywhere (the reference built with -ffp-contract=off).
 *  - The matching forward path (dct3d.c:52-59, 70-79, 106-111, 144-148,
 *    442-472) so synthetic chunks can be encoded.
 *  - Cube-mode raycast (src/shaders/raycast.slang:505-1039 with R3D_MODE 0,
 *    common.slang:45-61), the mip pyramid (rounded 2x box,
 *    tools/lodpack/main.c:216-227) and the dilated 8^3 occupancy grid
 *    (src/vk/vkbackend.c:794-870).
 *  - The sheet-volume generator that defines the synthetic workloads.
 */
#ifndef R3D_ORACLE_H
#define R3D_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- 3ddct decode ------------------------------------------------------ */

/* Per-chunk header, the decoded preamble of dct3d.c:550-554. */
typedef struct r3do_chunk_hdr {
  float vmin;
  float vspan;
  int32_t dcq; /* dc = dcq / 16 */
  uint32_t table; /* step-table index (unused by the single-chunk calls) */
} r3do_chunk_hdr;

/* cm[k*16+n] = (float)(c_k * cos(pi*(2n+1)*k/32)) in double (dct3d.c:38-49). */
void r3do_dct_basis(float cm[256]);
/* 46-entry step table of dct3d.c:131-143: step[f] = qstep * powf(1+f, 0.65). */
void r3do_step_table_qstep(float qstep, float step46[46]);
/* The configured table q(i,j,k) = 1 + 2(i+j+k), indexed by f = i+j+k. */
void r3do_step_table_linear(float step46[46]);

/* Decode one chunk from its levels (raster (z*16+y)*16+x). Corrections are
 * applied in order as norm[pos] += delta (dct3d.c:572-587, delta already
 * signed and multiplied by cq). pre_round (optional) receives the FP32 values
 * before roundf. Returns 1, or 0 when the header is invalid (dct3d.c:555-556);
 * on 0 nothing is written. */
int r3do_decode_chunk(const int16_t lvl[4096], const r3do_chunk_hdr *h, const float step46[46],
                      const uint16_t *corr_pos, const float *corr_delta, uint32_t ncorr,
                      uint8_t out[4096], float *pre_round);
/* The FP32 pipeline pieces, for golden-vector checks. */
void r3do_dequant(const int16_t lvl[4096], const float step46[46], float coef[4096]);
void r3do_idct3(const float coef[4096], float blk[4096]);
void r3do_fdct3(const float blk[4096], float coef[4096]);

/* Batch decode: chunk i reads levels + i*4096 and writes out + i*4096.
 * Corrections are CSR (corr_off[n+1]) or NULL. Runs on nthreads threads.
 * status[i] = 1/0 as r3do_decode_chunk. */
void r3do_decode_batch(uint32_t n, const int16_t *levels, const r3do_chunk_hdr *hdr,
                       const float *tables, uint32_t ntables, const uint32_t *corr_off,
                       const uint16_t *corr_pos, const float *corr_delta, uint8_t *out,
                       int32_t *status, int nthreads);

/* Encode one u8 chunk with a given step table into int16 levels, following
 * dct3d.c ENC_BODY + encode_float up to the quantiser (no entropy coding).
 * Levels are clipped to int16. */
void r3do_encode_chunk(const uint8_t in[4096], const float step46[46], int16_t lvl[4096],
                       r3do_chunk_hdr *h);
/* Encode every 16^3 chunk of an n^3-multiple-of-16 volume, chunk order
 * (cz*CY + cy)*CX + cx. */
void r3do_encode_volume(const uint8_t *vol, uint32_t nx, uint32_t ny, uint32_t nz,
                        const float step46[46], int16_t *levels, r3do_chunk_hdr *hdr,
                        int nthreads);

/* ---- volume preprocessing --------------------------------------------- */

/* Level l+1 from level l: (sum of the 2x2x2 box + 4) / 8, source index
 * clamped to the level's last voxel. dims of level l+1 = max(d >> 1, 1). */
void r3do_downsample(const uint8_t *src, uint32_t nx, uint32_t ny, uint32_t nz, uint8_t *dst);
/* Per-8^3-block max (partial edge blocks), then 3^3 max dilation. */
void r3do_occupancy(const uint8_t *vol, uint32_t nx, uint32_t ny, uint32_t nz, uint8_t *occ);

/* ---- raycast ------------------------------------------------------------ */

/* Byte-for-byte mirror of r3d_frame_params (src/render/render_types.h:47-94). */
typedef struct r3do_frame_params {
  float cam_origin[3]; float step_voxels;
  float cam_right[3];  float density;
  float cam_up[3];     float lod_bias;
  float cam_forward[3]; float max_mip;
  uint32_t viewport[2];
  uint32_t mode;
  uint32_t frame_index;
  float threshold;
  uint32_t slab_grid, slab_wz;
  float slab_z0, slab_nx, slab_ny, slab_px, slab_py;
  uint32_t slab_depth;
  uint32_t clip_valid;
  float clip_orig[6][2];
  float vol_r0[3]; float vol_tx;
  float vol_r1[3]; float vol_ty;
  float vol_r2[3]; float vol_tz;
  uint32_t brick_mode;
  float skip_gate;
  float slab_x0, slab_y0;
  uint32_t view_org;
  uint32_t view_flags;
  float overlay_gain;
  uint32_t overlay_flags;
} r3do_frame_params;

/* A mip pyramid as separate linear [z][y][x] levels. */
typedef struct r3do_volume {
  uint32_t nlev;
  uint32_t dims[16][3];
  const uint8_t *lev[16];
  uint32_t odim[3];
  const uint8_t *occ;
  /* optional: one bit per 16^3 brick of every level (level by level, x
   * fastest), set for each trilinear footprint the raycast reads (the brick
   * holding its first voxel); NULL = not tracked */
  uint32_t *touched;
} r3do_volume;

typedef struct r3do_counters {
  uint64_t samples; /* loop iterations (each fetches one volume sample) */
  uint64_t skips;   /* iterations that took the empty-space jump */
  uint64_t shaded;  /* samples that ran the gradient taps */
  uint64_t rays;    /* rays marched */
  uint64_t blended;        /* samples that blended two mip levels */
  uint64_t blended_shaded; /* shaded samples among them */
} r3do_counters;

/* quality: 0 = the six-tap central-difference gradient, 1 = the four-tap
 * tetrahedral gradient (R3D_QUALITY_FULL / _FAST, render_types.h:35-38,
 * selected per renderer by r3d_set_quality, render.h:241-242).
 * Render views into an RGBA FP32 image of img_w x img_h (each view at its
 * view_org). tf is 256x4 floats. Thread rows ty with ty % row_step == row0 are
 * rendered (row_step 1 = all). px_steps (optional, img_w*img_h) receives the
 * per-pixel loop-iteration count. Returns 0, or -1 for an unsupported view. */
int r3do_raycast(const r3do_volume *vol, const float *tf, const r3do_frame_params *views,
                 uint32_t nviews, float alpha_cutoff, float *img, uint32_t img_w, uint32_t img_h,
                 uint32_t *px_steps, uint32_t row0, uint32_t row_step, int nthreads,
                 r3do_counters *cnt, int quality);

/* Bricks-LOD mode (raycast.slang R3D_MODE 3 with the LOD flag): the
 * reference's page table (64-word header, per-level entries) and 128^3 slots,
 * slot s at atlas + s * 2 MiB. Views must carry brick_mode = abpa | 0x20000. */
typedef struct r3do_bricks {
  const uint32_t *page;
  uint32_t npage;
  const uint8_t *atlas;
  uint32_t abpa;
} r3do_bricks;
int r3do_raycast_bricks(const r3do_bricks *B, const float *tf, const r3do_frame_params *views, uint32_t nviews,
                        float alpha_cutoff, float *img, uint32_t img_w, uint32_t img_h, uint32_t *px_steps,
                        uint32_t row0, uint32_t row_step, int nthreads, r3do_counters *cnt, int quality);
/* Page table + atlas of a pyramid; see oracle_raycast.c. Returns page words or 0. */
uint32_t r3do_bricks_build(const r3do_volume *V, const uint8_t *resident, uint32_t *page, uint32_t page_cap,
                           uint8_t *atlas, uint32_t nslots);

/* Differentiable raycast in FP64 (parameter-dependent parts): renders
 * img (double RGBA, alpha = accumulated opacity) and, with dimg != NULL,
 * adds dL/dtf (256x4) and dL/dgrid. FULL or FLAT views only. */
int r3do_raycast_diff(const r3do_volume *vol, const double *tf, const double *grid, const uint32_t gd[3],
                      const r3do_frame_params *views, uint32_t nviews, double cutoff, double *img,
                      uint32_t img_w, uint32_t img_h, const double *dimg, double *dtf, double *dgrid);
/* The same, over the ray grid window [x0, x1) x [y0, y1) of every view only. */
int r3do_raycast_diff_window(const r3do_volume *vol, const double *tf, const double *grid, const uint32_t gd[3],
                             const r3do_frame_params *views, uint32_t nviews, double cutoff, double *img,
                             uint32_t img_w, uint32_t img_h, const double *dimg, double *dtf, double *dgrid,
                             uint32_t x0, uint32_t y0, uint32_t x1, uint32_t y1);

/* Planted mistakes (test sensitivity checks only; see oracle_raycast.c). */
enum {
  R3DO_PLANT_NONE = 0,
  R3DO_PLANT_SPEC_EXP = 1,     /* specular |n.h|^16 instead of ^24 */
  R3DO_PLANT_LIT_LERP = 2,     /* lerp(lit, 1, w) instead of lerp(1, lit, w) */
  R3DO_PLANT_TAP_RADIUS = 3,   /* cube taps at one voxel, not voxel * lod_step */
  R3DO_PLANT_MIP_WEIGHT = 4,   /* mip blend weight 1 - w */
  R3DO_PLANT_SKIP_NUDGE = 5,   /* no 1% nudge past the block face */
  R3DO_PLANT_SKIP_FACE = 6,    /* jump to the block's lower face, not the next one */
  R3DO_PLANT_FAST_RENORM = 7,  /* four-tap gradient without the sqrt(3)/2 renorm */
  R3DO_PLANT_KEY_LIGHT = 8,    /* diffuse without the key light term */
};
void r3do_plant(int k);

/* Pinned math, exported so tests can compare against libm and the GPU. */
float r3do_log2f(float x);
float r3do_exp2f(float x);
float r3do_pow_unit(float b, float y);
float r3do_hash12(uint32_t px, uint32_t py, uint32_t frame);
void r3do_math_batch(const float *x, uint32_t n, uint32_t fn, float *out);

/* ---- synthetic sheet volume ------------------------------------------ */

/* Concentric wavy cylindrical sheets about the z axis through (n/2, n/2):
 * sheet k has radius 20 + 9k + 3 sin(2 pi z / 40), is 3 voxels thick and has
 * value 160 + off_k (off_k uniform in [-30,30]); background 20 + U[-5,5];
 * plus N(0, 8) noise (Box-Muller); rounded, clamped to [0,255]. Writes voxels
 * [z0, z1) of an nx x ny x nz volume (x fastest). */
void r3do_synth_sheets(uint8_t *vol, uint32_t nx, uint32_t ny, uint32_t nz, uint32_t nsheets,
                       uint64_t seed, uint32_t z0, uint32_t z1, int nthreads);

/* ---- 3ddct entropy stage (oracle_entropy.c) ---- */
typedef struct r3do_blob_hdr {
  float vmin, vspan, qstep;
  int32_t dcq;
} r3do_blob_hdr;
int r3do_parse_blob(const uint8_t *blob, size_t len, uint32_t want_dtype, r3do_blob_hdr *h, int16_t lvl[4096],
                    uint32_t *ncorr, uint16_t *corr_pos, float *corr_delta);
size_t r3do_encode_blob(const r3do_blob_hdr *h, const int16_t lvl[4096], uint32_t dtype, uint8_t *out, size_t cap);
/* Blob batch decode (parse, qstep table, decode) as r3d_shard_decode_region
 * does per chunk; status[i] as r3do_parse_blob, zero-filled unless 1. */
void r3do_decode_blobs(uint32_t n, const uint8_t *bytes, const uint64_t *offsets, const uint32_t *lengths,
                       uint32_t dtype, uint8_t *out, int32_t *status, int nthreads);
/* Wall seconds of `iters` passes of decode (dct3d_decode_u8's signature) over
 * n blobs on nthreads threads, the loop of tools/3ddct/perf.c:55-60. */
double r3do_time_decoder(int (*decode)(const uint8_t *, size_t, uint8_t *), const uint8_t *bytes,
                         const uint64_t *offsets, const uint32_t *lengths, uint32_t n, uint32_t iters, int nthreads,
                         uint64_t *ok, uint64_t *sink);

/* ---- zarr v3 shard container (src/shard/shardio.c) ------------------- */

/* A present shard: the whole file in memory and its shard coordinates. */
typedef struct r3do_shard {
  const uint8_t *bytes;
  uint64_t n;
  uint32_t sz, sy, sx;
  uint32_t reserved0;
} r3do_shard;

/* CRC-32C (reflected 0x82F63B78, init and xorout 0xFFFFFFFF). */
uint32_t r3do_crc32c(const uint8_t *p, size_t n);
/* 1 when the shard is long enough for its index and the index crc matches. */
int r3do_shard_valid(const uint8_t *shard, uint64_t n);
/* 0 present, 1 missing, -1 invalid (shardio.c:157-172). */
int r3do_shard_chunk_location(const uint8_t *shard, uint64_t n, uint32_t cz, uint32_t cy, uint32_t cx,
                              uint64_t *off, uint64_t *len);
/* r3d_shard_decode_region over the listed shards; see oracle_shard.c. */
int r3do_shard_decode_region(const r3do_shard *shards, uint32_t nshards, uint64_t nz, uint64_t ny, uint64_t nx,
                             uint64_t z0, uint64_t y0, uint64_t x0, uint32_t dz, uint32_t dy, uint32_t dx,
                             uint8_t *dst, uint32_t counts[4], int32_t *chunk_status, int nthreads);

/* ---- decode-side deblocking, u8 (tools/3ddct/deblock.c) --------------- */

/* dct3d_deblock_chunk_u8: nbr = xlo, xhi, ylo, yhi, zlo, zhi (NULL: none). */
void r3do_deblock_chunk_u8(const uint8_t *chunk, const uint8_t *const nbr[6], float step, float strength,
                           uint8_t *out);
/* dct3d_deblock_u8 in place; dims z, y, x. */
void r3do_deblock_u8(uint8_t *vol, uint64_t nz, uint64_t ny, uint64_t nx, float step, float strength);

#ifdef __cplusplus
}
#endif
#endif /* R3D_ORACLE_H */
