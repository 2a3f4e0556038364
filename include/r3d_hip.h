/* r3d_hip.h -- C ABI of the MI355X (gfx950) path for render3d's two
 * data-parallel hot paths: the 3ddct 16^3 chunk decoder and the cube-mode
 * compute raycaster (plus its adjoint).
 *
 * Plain C: no HIP, torch or C++ types cross this boundary. Streams are passed
 * as `void *` holding a hipStream_t (NULL = the device's null stream).
 * Conventions:
 *   - Every int-returning call returns R3D_OK (0) or a negative r3d_status.
 *     (The reference's dct3d returns 1/0, tools/3ddct/dct3d.h:68-70; that
 *     convention is not carried across this ABI.) r3d_hip_last_error() gives
 *     the message of the last failure on a context.
 *   - Nothing is launched for an invalid call: bad counts, sizes beyond the
 *     stated limits, NULL required pointers, unsupported view kinds.
 *   - Bulk data (levels, pairs, headers, outputs) are device pointers owned by
 *     the caller; volumes and transfer tables are owned by their handles.
 *   - A context is used by one host thread at a time (or under the caller's
 *     lock), like the reference's single queue behind a mutex
 *     (src/vk/vkctx.c:329-343).
 */
#ifndef R3D_HIP_H
#define R3D_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status -------------------------------------------------------------- */
enum {
  R3D_OK = 0,
  R3D_ERR_INVALID = -1,     /* bad argument or malformed batch; nothing launched */
  R3D_ERR_NOMEM = -2,       /* device or host allocation failed */
  R3D_ERR_DEVICE = -3,      /* a HIP runtime call failed */
  R3D_ERR_UNSUPPORTED = -4, /* a view kind / variant the library does not implement */
  R3D_ERR_RANGE = -5,       /* a size beyond a documented limit */
};

typedef struct r3d_hip_ctx r3d_hip_ctx;
typedef struct r3d_hip_volume r3d_hip_volume;

/* ---- shared POD types (byte-for-byte mirrors of the reference) ------------ */

/* render3d's own src/render/render_types.h (include guard R3D_RENDER_TYPES_H)
 * defines the same PODs and mode values; when it was included first, its
 * definitions are used and these are skipped. */
#ifndef R3D_RENDER_TYPES_H

/* src/render/render_types.h:40-45 */
typedef struct r3d_volume_desc {
  uint32_t nx, ny, nz;
  uint32_t brick_dim;
  float voxel_um;
} r3d_volume_desc;

/* src/render/render_types.h:47-94 (248 bytes, mirrors FrameParams in
 * src/shaders/common.slang:5-42). Only cube-mode views are rendered: a view
 * with slab_grid, clip_valid or brick_mode nonzero, or the surf flag, is
 * rejected with R3D_ERR_UNSUPPORTED. */
typedef struct r3d_frame_params {
  float cam_origin[3];  float step_voxels;
  float cam_right[3];   float density;
  float cam_up[3];      float lod_bias;
  float cam_forward[3]; float max_mip;
  uint32_t viewport[2];
  uint32_t mode;
  uint32_t frame_index;
  float threshold;
  uint32_t slab_grid;
  uint32_t slab_wz;
  float slab_z0;
  float slab_nx, slab_ny;
  float slab_px, slab_py;
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
} r3d_frame_params;

/* render modes, render_types.h:25-33 */
enum {
  R3D_MODE_FULL = 0,
  R3D_MODE_MIP = 1,
  R3D_MODE_DEPTH = 2,
  R3D_MODE_HEATMAP = 3,
  R3D_MODE_RAYDIR = 4,
  R3D_MODE_FLAT = 5,
};
#define R3D_VIEW_ORTHO 1u
#define R3D_VIEW_HALF 32u

/* shading quality, render_types.h:35-38 */
enum {
  R3D_QUALITY_FULL = 0, /* six-tap central-difference gradient */
  R3D_QUALITY_FAST = 1, /* four-tap tetrahedral gradient */
};

#endif /* R3D_RENDER_TYPES_H */

/* ---- context ------------------------------------------------------------- */

/* Replaces r3d_create (src/render/render.h:17) for the compute path: binds a
 * HIP device and builds the lazily published constant tables (the DCT basis
 * of tools/3ddct/dct3d.c:38-49). */
int r3d_hip_init(r3d_hip_ctx **out, int device);
/* r3d_destroy (render.h:18). */
void r3d_hip_destroy(r3d_hip_ctx *ctx);
const char *r3d_hip_last_error(const r3d_hip_ctx *ctx);
/* Waits for `stream`. */
int r3d_hip_sync(r3d_hip_ctx *ctx, void *stream);
/* The 16x16 FP32 DCT basis the decoder uses, cm[k*16+n] (dct3d.c:38-49). */
int r3d_hip_dct_basis(const r3d_hip_ctx *ctx, float out[256]);

/* ---- 3ddct chunk decode ------------------------------------------------------ */

#define R3D_CHUNK_VOXELS 4096u
#define R3D_STEP_TABLE_LEN 46u        /* one step per L1 frequency 0..45 (dct3d.c:131-143) */
#define R3D_MAX_CORRECTIONS 4096u     /* per chunk, as dct3d.c:575 clamps */

/* Per-chunk header: the decoded preamble of dct3d.c:550-554. */
typedef struct r3d_chunk_header {
  float vmin;
  float vspan;     /* > 0 */
  int32_t dcq;     /* block mean in 1/16 units: dc = dcq / 16 */
  uint32_t table;  /* index into step_tables */
} r3d_chunk_header;

enum {
  R3D_COEF_DENSE_I16 = 0, /* levels[i*4096 + (z*16+y)*16+x], int16 */
  R3D_COEF_SPARSE = 1,    /* pairs[pair_offsets[i] .. pair_offsets[i+1]) */
};
/* A sparse pair: bits 0..11 = raster position, bits 16..31 = int16 level.
 * Each position appears at most once per chunk. */
#define R3D_PAIR(pos, level) ((uint32_t)(pos) | ((uint32_t)(uint16_t)(int16_t)(level) << 16))

/* One batch of chunks. All pointers are device pointers.
 * Output chunk i: voxel (z,y,x) goes to
 *   out + (out_offsets ? out_offsets[i] : i*4096) + z*out_stride_z + y*out_stride_y + x
 * (strides 0 -> 256 and 16: flat 4096-byte chunks). */
typedef struct r3d_chunk_batch {
  uint32_t nchunks;
  uint32_t coef_format;
  const int16_t *levels;          /* dense */
  const uint32_t *pair_offsets;   /* sparse: nchunks + 1 entries */
  const uint32_t *pairs;          /* sparse */
  uint64_t npairs;                /* sparse: entries in pairs; offsets are clamped to it */
  const r3d_chunk_header *headers;
  const float *step_tables;       /* ntables * 46 floats */
  uint32_t ntables;
  uint32_t reserved0;
  /* optional sparse corrections (dct3d.c:572-587): per chunk, applied in
   * order as norm[pos] += delta; delta carries the sign and cq. */
  const uint32_t *corr_offsets;   /* nchunks + 1, or NULL */
  const uint16_t *corr_pos;
  const float *corr_delta;
  uint64_t ncorr;                 /* entries in corr_pos / corr_delta */
  uint8_t *out;
  const uint64_t *out_offsets;    /* or NULL */
  uint32_t out_stride_y;
  uint32_t out_stride_z;
  /* optional per-chunk status: 1 = decoded, 0 = invalid header (the chunk is
   * zero-filled, as r3d_shard_chunk_decode does for a corrupt chunk,
   * src/shard/shardio.c:170-189). */
  int32_t *status;
  /* optional per-chunk counts for capacity layouts (r3d_hip_parse_blobs):
   * chunk c's pairs are pairs[pair_offsets[c] .. + pair_counts[c]) and its
   * corrections corr_pos/corr_delta[corr_offsets[c] .. + corr_counts[c]),
   * both still clamped to the next offset and to npairs / ncorr. */
  const uint32_t *pair_counts;
  const uint32_t *corr_counts;
} r3d_chunk_batch;

enum {
  R3D_DECODE_EXACT_F32 = 0,   /* the pinned FP32 sequence: bit-exact with the oracle */
  R3D_DECODE_MFMA_F32 = 1,    /* FP32 MFMA (16x16x4) for two passes, FP32 third pass: approximate */
  R3D_DECODE_MFMA_F16 = 2,    /* FP16 MFMA, FP32 accumulate: approximate */
  R3D_DECODE_MFMA_F16X3 = 3,  /* FP16 MFMA on hi/lo splits (3 products), FP32 accumulate:
                                 about FP32 accuracy, not bit-exact */
};
/* flags */
#define R3D_DECODE_F32_OUT 1u  /* write the FP32 pre-round values (dct3d.c:590-591) instead
                                  of u8: out is float[nchunks * 4096], flat (no offsets or
                                  strides) */
typedef struct r3d_decode_opts {
  uint32_t variant;
  uint32_t flags;
} r3d_decode_opts;

/* Decodes every chunk of the batch: dequantise (deq_one, dct3d.c:149-154),
 * inverse 3-D DCT (dct3d.c:113-118), + dc, corrections, denormalise
 * (dct3d.c:589-591), roundf and clamp to u8 (dct3d.c:625-639). The batched
 * replacement of dct3d_decode_u8 (dct3d.h:80-91) behind
 * r3d_shard_decode_region (src/shard/shardio.c:269-306) and of the GPU batch
 * seam r3d_vkc5d_decode (src/vk/vkc5d.h:26-32). Asynchronous on `stream`. */
int r3d_hip_decode_chunks(r3d_hip_ctx *ctx, const r3d_chunk_batch *batch,
                          const r3d_decode_opts *opts, void *stream);

/* ---- volumes --------------------------------------------------------------- */

#define R3D_MAX_LEVELS 16u
enum { R3D_SRC_HOST = 0, R3D_SRC_DEVICE = 1 };

/* Allocates a bricked mip pyramid (16^3 bricks with a one-voxel apron per
 * level) plus the dilated 8^3 occupancy grid. nlevels = 0 selects the full
 * chain floor(log2(maxdim)) + 1 (src/vk/vkbackend.c:1517-1518). */
int r3d_hip_volume_create(r3d_hip_ctx *ctx, const r3d_volume_desc *desc, uint32_t nlevels,
                          r3d_hip_volume **out);
void r3d_hip_volume_destroy(r3d_hip_volume *vol);
/* Level count and dims (dims[l*3 + {0,1,2}] = x, y, z) and occupancy dims. */
int r3d_hip_volume_info(const r3d_hip_volume *vol, uint32_t *nlevels, uint32_t dims[48],
                        uint32_t odim[3]);
/* r3d_upload_volume (render.h:20): copies linear [z][y][x] voxels (host or
 * device) into level 0, then builds aprons, the pyramid and the occupancy. */
int r3d_hip_volume_upload(r3d_hip_ctx *ctx, r3d_hip_volume *vol, const uint8_t *voxels,
                          uint32_t src_kind, void *stream);
/* Where chunk (cx,cy,cz) of level 0 lives, for decoding straight into the
 * volume: writes one byte offset per 16^3 chunk in (cz*CY + cy)*CX + cx order
 * into out_offsets (device, CX*CY*CZ entries) and returns the base pointer and
 * strides to put in an r3d_chunk_batch. */
int r3d_hip_volume_chunk_targets(r3d_hip_ctx *ctx, r3d_hip_volume *vol, uint64_t *out_offsets,
                                 uint8_t **base, uint32_t *stride_y, uint32_t *stride_z,
                                 void *stream);
/* Rebuilds level-0 aprons, levels 1.. (rounded 2x box) and the occupancy from
 * level 0 (after a decode into it through r3d_hip_volume_chunk_targets). */
int r3d_hip_volume_build(r3d_hip_ctx *ctx, r3d_hip_volume *vol, void *stream);
/* Builds the whole volume (level 0 with aprons, levels 1.., occupancy) from
 * flat decoded chunks (device, 4096 bytes each, (cz*CY+cy)*CX+cx order, 16-byte
 * aligned); level-0 dims must be multiples of 16. */
int r3d_hip_volume_build_from_chunks(r3d_hip_ctx *ctx, r3d_hip_volume *vol, const uint8_t *chunks,
                                     void *stream);
/* The shard path (r3d_shard_decode_region, src/shard/shardio.c:269-306, moved
 * onto the GPU): decodes every chunk of the volume ((cz*CY+cy)*CX+cx order;
 * the batch's out, out_offsets and strides are ignored) into a staging area the
 * volume owns, then builds it as r3d_hip_volume_build_from_chunks. */
int r3d_hip_volume_decode(r3d_hip_ctx *ctx, r3d_hip_volume *vol, const r3d_chunk_batch *batch,
                          const r3d_decode_opts *opts, void *stream);
/* Bricks over all levels of the pyramid (the bit count of a touched-brick
 * bitmap, r3d_raycast_opts.touched); each brick is R3D_HIP_BRICK_BYTES in HBM
 * (16^3 voxels, a one-voxel apron on the +x/+y/+z faces, padded to 16 B). */
#define R3D_HIP_BRICK_BYTES 4928u
uint64_t r3d_hip_volume_brick_count(const r3d_hip_volume *vol);
/* Copies level `level` out as linear [z][y][x] voxels (host or device dst). */
int r3d_hip_volume_read_level(r3d_hip_ctx *ctx, const r3d_hip_volume *vol, uint32_t level,
                              uint8_t *dst, uint32_t dst_kind, void *stream);
int r3d_hip_volume_read_occupancy(r3d_hip_ctx *ctx, const r3d_hip_volume *vol, uint8_t *dst,
                                  uint32_t dst_kind, void *stream);

/* ---- transfer function ------------------------------------------------------ */

/* r3d_set_transfer (render.h:240): 256 RGBA8 entries, stored as unorm floats. */
int r3d_hip_set_transfer(r3d_hip_ctx *ctx, const uint8_t rgba[256][4]);
/* FP32 table (256 x 4, host memory), for the differentiable path. */
int r3d_hip_set_transfer_f32(r3d_hip_ctx *ctx, const float *rgba);

/* ---- raycast --------------------------------------------------------------- */

/* Output image (device memory). rgba: FP32 RGBA, width*height*4 floats;
 * rgba8: RGBA8 unorm (round(sat(c)*255)). Either may be NULL. */
typedef struct r3d_hip_image {
  float *rgba;
  uint8_t *rgba8;
  uint32_t width, height;
} r3d_hip_image;

typedef struct r3d_hip_stats {
  uint64_t samples; /* loop iterations, each fetching one volume sample */
  uint64_t skips;   /* iterations that took the empty-space jump */
  uint64_t shaded;  /* samples that ran the gradient taps */
  uint64_t rays;
  /* with R3D_RAYCAST_MEASURE only (0 otherwise): */
  uint64_t blended;        /* samples that blended two mip levels (two trilinear footprints) */
  uint64_t blended_shaded; /* shaded samples among them (their taps blend too) */
  /* r3d_frame_stats (render_types.h:107-119): gpu_raycast_ns is the GPU time
   * of this call's launches (HIP events on the call's stream), panes_drawn
   * the views it rendered (pane-cache misses). The counters above cover the
   * drawn views only. */
  uint64_t gpu_raycast_ns;
  uint32_t panes_drawn;
  uint32_t reserved;
} r3d_hip_stats;

#define R3D_MAX_VIEWS_PER_CALL 256u
/* flags: measurement mode (r3d_hip_raycast only; no counterpart in the
 * reference). Slower 8x8 instances render the same image and also count
 * blended samples: counters then holds six u64 (samples, skips, shaded,
 * rays, blended, blended_shaded, as r3d_hip_stats) and touched, when given,
 * gets the bricks the footprints read. Not for timing. */
#define R3D_RAYCAST_MEASURE 1u
/* flags: the pane cache (vkbackend.c:7066-7079, 7144-7159). A view is drawn
 * only when its parameters (frame_index aside: a cached view keeps its last
 * jitter), the layout of all views (view_org, viewport), the outputs, the
 * options or the scene changed since view index v was last drawn with this
 * flag; the others keep the pixels already in the output. The scene changes
 * with every volume build or decode, bricks build or streamed fill, transfer
 * function, quality or tuning set; call r3d_hip_invalidate_panes after writes
 * the library does not see (the contents of opts->tf, the bricks tables
 * written directly, the output written by anything else). Any render without
 * the flag (raycast, plan launch, diff forward) makes every view redraw. */
#define R3D_RAYCAST_PANE_CACHE 2u
typedef struct r3d_raycast_opts {
  float alpha_cutoff;     /* early termination: raycast.slang:124 uses 0.98 */
  uint32_t flags;
  const float *tf;        /* optional device FP32 LUT (256x4); NULL = the context's */
  uint64_t *counters;     /* optional device u64[4] (u64[6] with R3D_RAYCAST_MEASURE):
                             stats accumulate here */
  uint32_t *px_steps;     /* optional device u32[width*height]: per-pixel iterations */
  /* optional, with R3D_RAYCAST_MEASURE: a device bitmap of
   * (r3d_hip_volume_brick_count + 31) / 32 words, zeroed by the caller;
   * every brick a sample's or a gradient tap's footprint reads gets its bit
   * (brick index: level by level, x fastest within a level). */
  uint32_t *touched;
} r3d_raycast_opts;

/* Renders `nviews` cube-mode views (each at its view_org) into `out`: the
 * compute dispatch of src/vk/vkbackend.c:7121-7173 running
 * src/shaders/raycast.slang main with R3D_MODE 0. `views` is host memory.
 * If `stats` is non-NULL the call waits for `stream` and fills it. */
int r3d_hip_raycast(r3d_hip_ctx *ctx, const r3d_hip_volume *vol, const r3d_frame_params *views,
                    uint32_t nviews, const r3d_hip_image *out, const r3d_raycast_opts *opts,
                    r3d_hip_stats *stats, void *stream);

/* r3d_set_quality (src/render/render.h:241-242): the gradient of FULL-mode
 * views, six central-difference taps (R3D_QUALITY_FULL, the default) or four
 * tetrahedral taps renormalised by sqrt(3)/2 (R3D_QUALITY_FAST,
 * src/shaders/raycast.slang:919-955), for later r3d_hip_raycast and
 * r3d_hip_bricks_raycast calls and for plans created after it (a plan keeps
 * the quality it was created with). Any other value selects FULL, as
 * src/vk/vkbackend.c:2075-2077 does. The differentiable path always uses six
 * taps. */
void r3d_hip_set_quality(r3d_hip_ctx *ctx, uint32_t quality);
/* The reference bumps its scene generation where its own data changes
 * (vkbackend.c, scene_gen++); here data the caller writes directly must be
 * announced: every pane-cached view redraws on its next call. */
void r3d_hip_invalidate_panes(r3d_hip_ctx *ctx);

/* Launch shape of the raycast kernel: the workgroup renders tile_w x
 * tile_h-pixel tiles (8x8, 16x4, 16x8 or 16x16); blocks_per_cu > 0 makes
 * the launch persistent, with that many workgroups per CU taking tiles from a
 * queue (at most 32); 0 launches one workgroup per tile. ms is filled by the
 * tuner. The reference dispatches 16x8 workgroups, one per tile
 * (src/vk/vkbackend.c:7121-7173). */
typedef struct r3d_raycast_tuning {
  uint32_t tile_w, tile_h;
  uint32_t blocks_per_cu;
  float ms;
} r3d_raycast_tuning;

/* Launch shape used by r3d_hip_raycast (default 8x8, one workgroup per tile). */
int r3d_hip_set_raycast_tuning(r3d_hip_ctx *ctx, const r3d_raycast_tuning *t);

/* A raycast plan holds a fixed view set and output on the device, so that a
 * launch is kernel launches only (no copies, allocations or host waits) and
 * can be captured into a hipGraph. No counterpart in the reference (its
 * Vulkan dispatch records a command buffer per frame). The plan keeps
 * pointers to vol, the image buffers, opts->tf, opts->counters (stats
 * accumulate there on every launch) and opts->px_steps; they must stay valid
 * while the plan is used. */
typedef struct r3d_hip_raycast_plan r3d_hip_raycast_plan;
int r3d_hip_raycast_plan_create(r3d_hip_ctx *ctx, const r3d_hip_volume *vol, const r3d_frame_params *views,
                                uint32_t nviews, const r3d_hip_image *out, const r3d_raycast_opts *opts,
                                r3d_hip_raycast_plan **plan);
int r3d_hip_raycast_plan_launch(r3d_hip_raycast_plan *plan, void *stream);
/* Autotuner: times each tile shape at one workgroup per tile and at 2, 4, 8
 * and 16 persistent workgroups per CU, keeps the fastest in the plan and
 * writes up to max_results measurements. Synchronises `stream`: call it
 * before capturing the plan, not during. */
int r3d_hip_raycast_plan_tune(r3d_hip_raycast_plan *plan, void *stream, r3d_raycast_tuning *results,
                              uint32_t max_results, uint32_t *nresults);
int r3d_hip_raycast_plan_set_tuning(r3d_hip_raycast_plan *plan, const r3d_raycast_tuning *t);
int r3d_hip_raycast_plan_get_tuning(const r3d_hip_raycast_plan *plan, r3d_raycast_tuning *t);
void r3d_hip_raycast_plan_destroy(r3d_hip_raycast_plan *plan);

/* ---- differentiable raycast -------------------------------------------------- */

/* No counterpart in the reference (it has no differentiable path). The
 * compositing integral of raycast.slang:893-1038 with an FP32 transfer table
 * and a per-voxel density-correction grid g (trilinear, cell-centred over the
 * unit cube) that scales the extinction: a = 1 - (1 - sat(s.a))^(density *
 * step * lod_step * g). Output alpha is the accumulated opacity. Empty-space
 * skipping is off. All views of one call are FULL (shaded) or all FLAT. */
typedef struct r3d_diff_params {
  const float *tf;          /* device, 256 x 4 FP32 */
  const float *grid;        /* device, grid_dims[2] x [1] x [0] FP32 */
  uint32_t grid_dims[3];    /* x, y, z */
  float alpha_cutoff;       /* 0 -> 0.98 */
  void *workspace;          /* device, >= r3d_hip_diff_workspace_bytes(views, nviews) */
  size_t workspace_bytes;
} r3d_diff_params;

/* Bytes of per-ray checkpoint state (loop state every 64 iterations). */
size_t r3d_hip_diff_workspace_bytes(const r3d_frame_params *views, uint32_t nviews);
/* Forward: renders into out->rgba and writes the checkpoints to the
 * workspace. With dp->workspace NULL it renders the same image and keeps no
 * checkpoints (a forward-only render of the differentiable model). */
int r3d_hip_raycast_diff(r3d_hip_ctx *ctx, const r3d_hip_volume *vol, const r3d_frame_params *views,
                         uint32_t nviews, const r3d_hip_image *out, const r3d_diff_params *dp,
                         void *stream);
/* Adjoint of the preceding r3d_hip_raycast_diff with the same arguments:
 * dL_drgba is the gradient of the loss with respect to that call's FP32
 * image (img_w x img_h x 4); writes dL_dtf (256 x 4) and ADDS into dL_dgrid
 * (grid-shaped; zero it first). Segments of 64 iterations are recomputed from
 * their checkpoints and run in reverse. */
int r3d_hip_raycast_adjoint(r3d_hip_ctx *ctx, const r3d_hip_volume *vol, const r3d_frame_params *views,
                            uint32_t nviews, const float *dL_drgba, uint32_t img_w, uint32_t img_h,
                            float *dL_dtf, float *dL_dgrid, const r3d_diff_params *dp, void *stream);

/* ---- synthetic workloads ------------------------------------------------------- */

/* The seeded sheet volume (linear [z][y][x], device dst): concentric wavy
 * cylindrical sheets about the z axis, noise N(0,8). Bit-identical to the
 * host generator of the same definition. */
int r3d_hip_synth_sheets(r3d_hip_ctx *ctx, uint8_t *dst, uint32_t nx, uint32_t ny, uint32_t nz,
                         uint32_t nsheets, uint64_t seed, void *stream);
/* Forward 3-D DCT + dead-zone quantiser of every 16^3 chunk of a linear
 * volume (dims multiples of 16), ENC_BODY + encode_float of
 * tools/3ddct/dct3d.c:431-472 and 598-621 without the entropy stage.
 * Writes int16 levels (chunk order (cz*CY+cy)*CX+cx) and headers (table 0). */
int r3d_hip_encode_chunks(r3d_hip_ctx *ctx, const uint8_t *vol, uint32_t nx, uint32_t ny,
                          uint32_t nz, const float *step_table, int16_t *levels,
                          r3d_chunk_header *headers, void *stream);

/* ---- 3ddct blobs: the entropy stage ------------------------------------------ */

/* A 3ddct blob (tools/3ddct/dct3d.c:414-433): the preamble [0xD3][dtype], then
 * one range-coded stream with vmin, vspan, qstep, dcq, the corrections flag,
 * the levels and the corrections. dtype ids as dct3d.c:430. */
enum { R3D_DT_U8 = 1, R3D_DT_U16 = 2, R3D_DT_U32 = 3, R3D_DT_S8 = 4, R3D_DT_S16 = 5, R3D_DT_S32 = 6,
       R3D_DT_F32 = 7 };

typedef struct r3d_blob_batch {
  uint32_t nchunks;
  uint32_t dtype;               /* the dtype every blob must carry (R3D_DT_U8 for u8 output) */
  uint32_t max_pairs;           /* per-chunk capacity of levels and of corrections, 1..4096 */
  uint32_t reserved0;
  const uint8_t *bytes;         /* device: the blobs */
  const uint64_t *offsets;      /* device [nchunks]: first byte of each blob */
  const uint32_t *lengths;      /* device [nchunks]: length of each blob */
} r3d_blob_batch;

/* Device outputs of r3d_hip_parse_blobs, laid out as a sparse
 * r3d_chunk_batch with per-chunk counts: chunk c's data start at
 * c * max_pairs. */
typedef struct r3d_blob_parse_out {
  r3d_chunk_header *headers;    /* [nchunks]: table = c; vspan = 0 for a rejected blob */
  float *step_tables;           /* [nchunks * 46]: qstep * powf(1 + f, 0.65f) (dct3d.c:131-143) */
  uint32_t *pair_offsets;       /* [nchunks + 1] */
  uint32_t *pair_counts;        /* [nchunks] */
  uint32_t *pairs;              /* [nchunks * max_pairs]: R3D_PAIR(pos, level) */
  uint32_t *corr_offsets;       /* [nchunks + 1] */
  uint32_t *corr_counts;        /* [nchunks] */
  uint16_t *corr_pos;           /* [nchunks * max_pairs] */
  float *corr_delta;            /* [nchunks * max_pairs]: +-mag * cq (dct3d.c:583-585) */
  int32_t *status;              /* [nchunks] or NULL: 1 parsed; 0 malformed (decode_float would
                                   return 0); -1 a level beyond int16; -2 more levels or
                                   corrections than max_pairs. A chunk with status != 1 decodes
                                   to zeros with decode status 0. */
} r3d_blob_parse_out;

/* Parses each blob (the entropy part of decode_float, dct3d.c:542-587), one
 * GPU lane per blob. The result feeds r3d_hip_decode_chunks as a sparse batch
 * (R3D_COEF_SPARSE, ntables = nchunks, with pair_counts / corr_counts), which
 * together give dct3d_decode_u8's output. Replaces the per-chunk CPU loop of
 * r3d_shard_decode_region (src/shard/shardio.c:269-306).
 * Batches of 512 blobs or more are parsed longest first: the lane order and
 * its radix sort live in a scratch buffer of the context (12 bytes per blob
 * plus the sort's own), allocated the first time a batch needs more than the
 * context holds. A larger batch later takes a new buffer and keeps the old
 * one until r3d_hip_destroy, so a HIP graph that captured a parse stays
 * valid. Growing is refused (R3D_ERR_INVALID, nothing launched) while the
 * stream is capturing: parse a batch at least as large before capture.
 * Calls on one context share that buffer, so they must not overlap in time
 * (one stream, or ordered by the caller). */
int r3d_hip_parse_blobs(r3d_hip_ctx *ctx, const r3d_blob_batch *in, const r3d_blob_parse_out *out, void *stream);

/* No counterpart in the reference: r3d_hip_parse_blobs for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc). The same parse and outputs, with
 * the lane order and its radix sort in a fixed device array of 16 MiB per
 * GPU (held until the process ends) instead of context memory: the call
 * allocates nothing. At most R3D_PARSE_NOALLOC_MAX_BLOBS blobs per call
 * (R3D_ERR_RANGE beyond, nothing launched). Every call on one GPU shares that
 * array, whatever its context: calls that overlap in time on one GPU (other
 * streams, contexts or threads) race. */
#define R3D_PARSE_NOALLOC_MAX_BLOBS 262144u
int r3d_hip_parse_blobs_noalloc(r3d_hip_ctx *ctx, const r3d_blob_batch *in, const r3d_blob_parse_out *out,
                                void *stream);

/* Workload encoder: the entropy-coding back half of encode_float
 * (dct3d.c:499-537) without corrections, one lane per chunk. Blob c is written
 * at out + c * slot_bytes and lengths[c] is its length (longer than
 * slot_bytes if it did not fit; the excess is dropped). headers give vmin,
 * vspan and dcq; qsteps[c] is coded as the blob's qstep. */
int r3d_hip_encode_blobs(r3d_hip_ctx *ctx, uint32_t nchunks, const int16_t *levels,
                         const r3d_chunk_header *headers, const float *qsteps, uint32_t dtype, uint8_t *out,
                         uint64_t slot_bytes, uint32_t *lengths, void *stream);

/* ---- zarr v3 shards of 3ddct chunks -------------------------------------
 * A shard is one 1024^3 region of the volume: its 16^3 chunks' blobs back to
 * back, then an index of one (u64le offset, u64le nbytes) pair per chunk in
 * z-major order, (~0, ~0) for a missing chunk (zeros), then the crc32c of the
 * index (tools/3ddct/tools/export/shard.h:1-18; src/shard/shardio.h:1-5).
 * Replaces the reader src/shard/shardio.c for region decodes. */
#define R3D_HIP_SHARD_GRID 64u
#define R3D_HIP_SHARD_INDEX_BYTES (64u * 64u * 64u * 16u + 4u)

/* One present shard, resident in device memory: the whole file. */
typedef struct r3d_shard_src {
  const uint8_t *bytes;         /* device */
  uint64_t nbytes;
  uint32_t sz, sy, sx;          /* shard coordinates, as in <sz>_<sy>_<sx>.shard */
  uint32_t reserved0;
} r3d_shard_src;

/* The store's volume dims (r3d_shard_store, shardio.h:16-20) and a region of
 * it (the arguments of r3d_shard_decode_region, shardio.h:59-63). */
typedef struct r3d_shard_region {
  uint64_t nz, ny, nx;
  uint64_t z0, y0, x0;
  uint32_t dz, dy, dx;
  uint32_t reserved0;
} r3d_shard_region;

/* Chunks whose voxels the region decode writes: the chunks that intersect
 * the region clamped to the volume, as a box (x fastest). The length of
 * chunk_status below; 0 for an empty region or one outside the volume. */
uint64_t r3d_hip_shard_region_chunks(const r3d_shard_region *region);

/* r3d_shard_decode_region (shardio.c:269-306) on the GPU. dst (device,
 * dz * dy * dx bytes, x fastest) receives the region; voxels outside the
 * volume, in absent shards (the shards not listed) or in missing chunks are
 * 0. Each listed shard is checked on the device (long enough for its index,
 * index crc32c) and every chunk's index entry is checked against the payload
 * (shardio.c:157-172). chunk_status (device, optional,
 * r3d_hip_shard_region_chunks entries in chunk-box order): 1 decoded, 0
 * missing or absent, -1 corrupt (a bad shard, index entry or blob: zeros),
 * -2 levels beyond int16 or over the 4096-entry capacity (zeros; the
 * reference decodes these: decode them with dct3d_decode_u8 on the host).
 * counts (device, optional, 4 x u32) accumulates decoded, missing, corrupt
 * and -2 chunks; the reference returns -1 exactly when counts[2] grows.
 * Decoded voxels are bit-identical to dct3d_decode_u8. The context keeps a
 * workspace of about 41 KB per chunk, up to 262,144 chunks at a time.
 * Returns 0, or a negative R3D_ERR_* for invalid arguments (a zero size, a
 * size that overflows) before any launch. */
int r3d_hip_shard_decode_region(r3d_hip_ctx *ctx, const r3d_shard_src *shards, uint32_t nshards,
                                const r3d_shard_region *region, uint8_t *dst, int32_t *chunk_status,
                                uint32_t *counts, void *stream);

/* The shard path straight into a render volume: r3d_hip_shard_decode_region
 * of region (whose size must be the volume's dims) into the volume's staging
 * buffer, then the bricks, aprons, pyramid and occupancy built from it, on
 * `stream` with no host round trip (the reference reads regions into host
 * memory and uploads them, shardio.c:269-306 then r3d_upload_volume). The
 * statuses and counts are those of the region decode. */
int r3d_hip_volume_decode_shards(r3d_hip_ctx *ctx, r3d_hip_volume *vol, const r3d_shard_src *shards,
                                 uint32_t nshards, const r3d_shard_region *region, int32_t *chunk_status,
                                 uint32_t *counts, void *stream);

/* CRC-32C (reflected Castagnoli polynomial, init and final xor 0xFFFFFFFF:
 * the zarr index trailer, shardio.c:39-52) of n device bytes into *crc
 * (device). */
int r3d_hip_crc32c(r3d_hip_ctx *ctx, const uint8_t *data, uint64_t n, uint32_t *crc, void *stream);

/* ---- decode-side deblocking (tools/3ddct/deblock.h), u8 ---------------- */

/* dct3d_deblock_u8 (deblock.c:132-203): in place over a dense device volume
 * (dims z, y, x; x fastest), every internal plane at a multiple of 16 voxels,
 * x then y then z seams. step is the encode quality (thresholds 2.0 / 1.2 x
 * step x strength); step <= 0 derives them from the data range over a coarse
 * subsample (deblock.c:144-152). strength <= 0 is a no-op. Bit-identical to
 * the reference. */
int r3d_hip_deblock_u8(r3d_hip_ctx *ctx, uint8_t *vol, const uint64_t dims[3], float step, float strength,
                       void *stream);

/* dct3d_deblock_chunk_u8 (deblock.c:69-130) for a batch: chunk c (device,
 * 4096 bytes each, z-major, 16-byte aligned) is filtered against its
 * pristine face neighbours neighbors[c * 6 + f] (device; f = xlo, xhi, ylo,
 * yhi, zlo, zhi; indices into chunks, negative or >= nchunks for none), faces
 * in that order, into out (must not overlap chunks). */
int r3d_hip_deblock_chunks_u8(r3d_hip_ctx *ctx, uint32_t nchunks, const uint8_t *chunks, const int32_t *neighbors,
                              float step, float strength, uint8_t *out, void *stream);

/* ---- bricks-LOD volumes (raycast.slang R3D_MODE 3 with BR_LOD_FLAG) ------
 * A volume as a pyramid of 128^3 bricks, each resident in a slot of an atlas
 * or absent, found through the reference's page table (vkbackend.c:4185-4205):
 * page[0] = nlev, page[1..3] = base shape (x, y, z), per level l at 4 + 4l
 * {offset of its entries, bx | by << 10 | bz << 20, scale = 2^l}; an entry is
 * slot | max << 24 (max = the brick's largest voxel), 0xFFFFFFFF when absent.
 * Slots are 128^3 u8, z-major, slot s at atlas + s * 128^3. A sample resolves
 * the desired level, then +1, +2 and the coarsest (raycast.slang:198-279).
 * Replaces the bricks-LOD dispatch of vkbackend.c for the render. */
typedef struct r3d_hip_bricks r3d_hip_bricks;

typedef struct r3d_bricks_desc {
  uint32_t nx, ny, nz;          /* base shape in voxels */
  uint32_t nlev;                /* levels, 1..15: level l has dims max(1, dim >> l) */
  uint32_t abpa;                /* atlas slots per axis (1..255): abpa^3 slots */
  uint32_t reserved0;
} r3d_bricks_desc;

/* Allocates the page table, the abpa^3-slot atlas and the build's scratch
 * (levels 1.. as linear arrays, about 1/7 of the base volume, and the job
 * list), so that a build allocates nothing. */
int r3d_hip_bricks_create(r3d_hip_ctx *ctx, const r3d_bricks_desc *desc, r3d_hip_bricks **out);
void r3d_hip_bricks_destroy(r3d_hip_bricks *b);
/* Builds the pyramid of vol (device, nz x ny x nx u8, x fastest) with the
 * rounded 2x box (tools/lodpack/main.c:216-227) and makes resident the bricks
 * whose resident[] flag is set (host, one byte per page entry in page order
 * after the 64-word header; NULL = all), in slots taken in page order; the
 * others read 0xFFFFFFFF. Voxels of a slot past its level's edge are 0.
 * Runs on the stream after its earlier work and returns when done. */
int r3d_hip_bricks_build(r3d_hip_ctx *ctx, r3d_hip_bricks *b, const uint8_t *vol, const uint8_t *resident,
                         void *stream);
/* The device page table and atlas, for a streamer that manages residency
 * itself (write slots, then entries, between frames). */
int r3d_hip_bricks_tables(const r3d_hip_bricks *b, uint32_t **page, uint32_t *page_words, uint8_t **atlas,
                          uint32_t *nslots);
/* r3d_frame_views for bricks-LOD views (brick_mode = abpa | 0x20000; zoom
 * crop, axis and oblique slabs as the reference; overlays off): as
 * r3d_hip_raycast otherwise. */
int r3d_hip_bricks_raycast(r3d_hip_ctx *ctx, const r3d_hip_bricks *b, const r3d_frame_params *views, uint32_t nviews,
                           const r3d_hip_image *out, const r3d_raycast_opts *opts, r3d_hip_stats *stats,
                           void *stream);

/* ---- streaming residency for bricks-LOD volumes ---------------------------
 * The bricks pump of the reference (src/render/render.h:115-132;
 * src/vk/vkbackend.c:5244-5599) over a handle's page table, for volumes whose
 * pyramid has more bricks than the atlas has slots. The warm tier is the
 * pyramid of the last r3d_hip_bricks_build, kept in HBM (level 0 is the
 * caller's volume: it must stay valid while streaming; the build's resident
 * flags seed the atlas, and the coarsest level should be resident, as the
 * reference makes it). Per frame: begin, then any number of box / point
 * collects, then submit; r3d_hip_bricks_stream is begin + the view-cone
 * collect + submit. Requested bricks that are resident are marked wanted;
 * the others become candidates, parent before child, then nearest first.
 * Bricks whose largest voxel is below the gate (at least 5/255, the codec
 * noise floor, vkbackend.c:2348) never take a slot. A submit fills at most
 * min(budget, 32) bricks into free slots, else into the least recently
 * wanted slots not wanted this frame (the evicted bricks' entries become
 * absent first), and publishes their entries (slot | max << 24). The fills
 * and entry writes are queued on `stream`, ordered after earlier frames on
 * it: use one stream for the renders and the stream calls (or order them).
 * Between builds the streamer owns the table: do not mix it with writes
 * through r3d_hip_bricks_tables. Geometry is in the renderer's volume space
 * (base voxels / the largest base dim). */

/* One level's 3ddct shard store: the level's shard files resident in device
 * memory (a shard not listed is absent), the store's volume being the level's
 * dims (x, y, z of the pyramid the handle was created for). */
typedef struct r3d_bricks_level_src {
  const r3d_shard_src *shards;
  uint32_t nshards;
  uint32_t reserved0;
} r3d_bricks_level_src;
/* The reference's warm tier in compressed form (its bricks come from a
 * pyramid of 3ddct stores): instead of a built pyramid, the source of the
 * fills is one shard store per level (nlev entries; the lists are copied, the
 * shard bytes must stay valid while streaming). Each fill decodes the
 * brick's 128^3 box of its level (r3d_hip_shard_decode_region) straight into
 * the slot; its maximum is measured on the GPU, the entry is published only
 * at 5 or more (the noise floor, vkbackend.c:5265-5269), and the gate learns
 * it at the next begin (a brick's maximum is unknown, so not gated, until it
 * has been decoded once). Bricks whose resident flag is set (one per page
 * entry, page order, or NULL for all) are filled now, in slots in page order;
 * the call returns when they are. While a submit's fills are in flight,
 * r3d_hip_bricks_stream_begin returns 0 (stats.inflight holds their count),
 * as the reference's pump skips a frame while its worker runs. */
int r3d_hip_bricks_build_from_shards(r3d_hip_ctx *ctx, r3d_hip_bricks *b, const r3d_bricks_level_src *levels,
                                     uint32_t nlev, const uint8_t *resident, void *stream);

/* r3d_bricks_stream_begin (render.h:127): opens a collect pass; returns 1, or
 * 0 when the handle has not been built (nothing is collected then), or while
 * the last submit's fills from shard stores are in flight. */
int r3d_hip_bricks_stream_begin(r3d_hip_bricks *b);
/* r3d_bricks_stream_box (render.h:128-129): the bricks of the box lo..hi at
 * the level whose voxel pitch matches pixel_cone (coarser until at most
 * 2,048 bricks), each with its parent. */
void r3d_hip_bricks_stream_box(r3d_hip_bricks *b, const float lo[3], const float hi[3], float pixel_cone,
                               float gate);
/* r3d_bricks_stream_point (render.h:130-131): the brick holding p at level,
 * and its parent. */
void r3d_hip_bricks_stream_point(r3d_hip_bricks *b, const float p[3], uint32_t level, float gate);
/* r3d_bricks_stream_submit (render.h:132): fills up to min(budget, 32)
 * candidates on `stream`. */
int r3d_hip_bricks_stream_submit(r3d_hip_ctx *ctx, r3d_hip_bricks *b, uint32_t budget, void *stream);
/* r3d_bricks_stream (render.h:115-124): begin, the view-cone collect (eye and
 * fwd in volume space, half_tan = tan(fov_y / 2) x the frustum diagonal
 * factor, pixel_cone = the ray-cone growth per unit distance; slice_depth > 0
 * limits z to [slice_z0, slice_z0 + slice_depth) base voxels), submit. */
int r3d_hip_bricks_stream(r3d_hip_ctx *ctx, r3d_hip_bricks *b, const float eye[3], const float fwd[3], float half_tan,
                          float pixel_cone, uint32_t slice_z0, uint32_t slice_depth, float gate, uint32_t budget,
                          void *stream);
/* r3d_bricks_get_stats (render.h:146-158), the fields a GPU-resident warm
 * tier has. */
typedef struct r3d_bricks_stream_stats {
  uint32_t nb, hot, hot_cap;  /* page entries; resident slots; slots */
  uint32_t inflight;          /* bricks filled by this pass's submit (0 from begin until it fills) */
  uint32_t evicted_last;      /* bricks evicted by this pass's submit */
  uint32_t nlevels;
  uint32_t lod_wanted[8];     /* the last pass's desired bricks per level */
  uint64_t lod_requests[8];   /* cumulative */
  uint64_t decoded, evicted, jobs; /* cumulative fills, evictions, submits that filled */
} r3d_bricks_stream_stats;
int r3d_hip_bricks_stream_stats(const r3d_hip_bricks *b, r3d_bricks_stream_stats *st);

#ifdef __cplusplus
}
#endif

#ifdef __cplusplus
static_assert(sizeof(r3d_frame_params) == 248, "r3d_frame_params must be 248 bytes");
static_assert(sizeof(r3d_chunk_header) == 16, "r3d_chunk_header must be 16 bytes");
#else
_Static_assert(sizeof(r3d_frame_params) == 248, "r3d_frame_params must be 248 bytes");
_Static_assert(sizeof(r3d_chunk_header) == 16, "r3d_chunk_header must be 16 bytes");
#endif

#endif /* R3D_HIP_H */
