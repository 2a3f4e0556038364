/* A C23 host program on the GPU path, through r3d_hip.h only (run by
 * tests/test_gpu_c23.py): it decodes the reference's golden 3ddct blobs
 * (parse, then the exact decode) and raycasts a small scene, writes what it
 * got, and compares the bytes with the expected files in <dir>.
 *
 *   <dir>/blobs.bin     the blobs back to back
 *   <dir>/offsets.bin   u64 per blob       <dir>/lengths.bin  u32 per blob
 *   <dir>/expect_u8.bin 4096 bytes per blob (dct3d_decode_u8's output)
 *   <dir>/scene.bin     u32 nx, ny, nz, width, height, nviews; nviews
 *                       r3d_frame_params; the RGBA8 transfer table (1024 B);
 *                       the volume (x fastest)
 *   <dir>/expect_rgba8.bin, expect_rgba.bin  the image (RGBA8, FP32)
 * Exit status 0 when everything matches (FP32 within 1e-5). */
#include <hip/hip_runtime_api.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "r3d_hip.h"

#define CHECK_HIP(x)                                                      \
  do {                                                                    \
    hipError_t e_ = (x);                                                  \
    if (e_ != hipSuccess) {                                               \
      fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));             \
      return 20;                                                          \
    }                                                                     \
  } while (0)
#define CHECK_R3D(x)                                                      \
  do {                                                                    \
    int r_ = (x);                                                         \
    if (r_ != R3D_OK) {                                                   \
      fprintf(stderr, "%s: %d (%s)\n", #x, r_, r3d_hip_last_error(ctx)); \
      return 21;                                                          \
    }                                                                     \
  } while (0)

static const char *g_dir;

static void *slurp(const char *name, size_t *n) {
  char path[1024];
  snprintf(path, sizeof path, "%s/%s", g_dir, name);
  FILE *f = fopen(path, "rb");
  if (!f) return NULL;
  fseek(f, 0, SEEK_END);
  const long len = ftell(f);
  fseek(f, 0, SEEK_SET);
  void *p = malloc(len > 0 ? (size_t)len : 1u);
  if (p && len > 0 && fread(p, 1, (size_t)len, f) != (size_t)len) {
    free(p);
    p = NULL;
  }
  fclose(f);
  *n = len > 0 ? (size_t)len : 0u;
  return p;
}

static int spit(const char *name, const void *p, size_t n) {
  char path[1024];
  snprintf(path, sizeof path, "%s/%s", g_dir, name);
  FILE *f = fopen(path, "wb");
  if (!f) return 1;
  const size_t w = fwrite(p, 1, n, f);
  fclose(f);
  return w != n;
}

static int decode_golden(r3d_hip_ctx *ctx) {
  size_t nb, no, nl, ne;
  uint8_t *bytes = slurp("blobs.bin", &nb);
  uint64_t *offs = slurp("offsets.bin", &no);
  uint32_t *lens = slurp("lengths.bin", &nl);
  uint8_t *expect = slurp("expect_u8.bin", &ne);
  if (!bytes || !offs || !lens || !expect) return 30;
  const uint32_t n = (uint32_t)(nl / 4u);
  if (no != 8u * n || ne != 4096u * n) return 31;
  const size_t cap = 4096u;
  uint8_t *d_bytes, *d_u8;
  uint64_t *d_offs;
  uint32_t *d_lens;
  r3d_blob_parse_out po;
  CHECK_HIP(hipMalloc((void **)&d_bytes, nb));
  CHECK_HIP(hipMalloc((void **)&d_offs, no));
  CHECK_HIP(hipMalloc((void **)&d_lens, nl));
  CHECK_HIP(hipMalloc((void **)&d_u8, 4096u * n));
  CHECK_HIP(hipMalloc((void **)&po.headers, sizeof(r3d_chunk_header) * n));
  CHECK_HIP(hipMalloc((void **)&po.step_tables, sizeof(float) * 46u * n));
  CHECK_HIP(hipMalloc((void **)&po.pair_offsets, 4u * (n + 1u)));
  CHECK_HIP(hipMalloc((void **)&po.pair_counts, 4u * n));
  CHECK_HIP(hipMalloc((void **)&po.pairs, 4u * cap * n));
  CHECK_HIP(hipMalloc((void **)&po.corr_offsets, 4u * (n + 1u)));
  CHECK_HIP(hipMalloc((void **)&po.corr_counts, 4u * n));
  CHECK_HIP(hipMalloc((void **)&po.corr_pos, 2u * cap * n));
  CHECK_HIP(hipMalloc((void **)&po.corr_delta, 4u * cap * n));
  CHECK_HIP(hipMalloc((void **)&po.status, 4u * n));
  CHECK_HIP(hipMemcpy(d_bytes, bytes, nb, hipMemcpyHostToDevice));
  CHECK_HIP(hipMemcpy(d_offs, offs, no, hipMemcpyHostToDevice));
  CHECK_HIP(hipMemcpy(d_lens, lens, nl, hipMemcpyHostToDevice));

  const r3d_blob_batch in = {.nchunks = n, .dtype = R3D_DT_U8, .max_pairs = (uint32_t)cap,
                             .bytes = d_bytes, .offsets = d_offs, .lengths = d_lens};
  CHECK_R3D(r3d_hip_parse_blobs(ctx, &in, &po, NULL));
  r3d_chunk_batch b;
  memset(&b, 0, sizeof b);
  b.nchunks = n;
  b.coef_format = R3D_COEF_SPARSE;
  b.pairs = po.pairs;
  b.pair_offsets = po.pair_offsets;
  b.pair_counts = po.pair_counts;
  b.npairs = cap * n;
  b.headers = po.headers;
  b.step_tables = po.step_tables;
  b.ntables = n;
  b.corr_offsets = po.corr_offsets;
  b.corr_counts = po.corr_counts;
  b.corr_pos = po.corr_pos;
  b.corr_delta = po.corr_delta;
  b.ncorr = cap * n;
  b.out = d_u8;
  const r3d_decode_opts o = {R3D_DECODE_EXACT_F32, 0u};
  CHECK_R3D(r3d_hip_decode_chunks(ctx, &b, &o, NULL));
  CHECK_HIP(hipDeviceSynchronize());

  uint8_t *got = malloc(4096u * n);
  int32_t *status = malloc(4u * n);
  if (!got || !status) return 32;
  CHECK_HIP(hipMemcpy(got, d_u8, 4096u * n, hipMemcpyDeviceToHost));
  CHECK_HIP(hipMemcpy(status, po.status, 4u * n, hipMemcpyDeviceToHost));
  uint32_t bad_status = 0, bad_chunks = 0;
  for (uint32_t i = 0; i < n; i++) {
    bad_status += status[i] != 1;
    bad_chunks += memcmp(got + 4096u * i, expect + 4096u * i, 4096u) != 0;
  }
  printf("decode: %u blobs, %u not parsed, %u differ\n", n, bad_status, bad_chunks);
  if (spit("got_u8.bin", got, 4096u * n)) return 33;
  void *dev[] = {d_bytes, d_offs, d_lens, d_u8, po.headers, po.step_tables, po.pair_offsets, po.pair_counts,
                 po.pairs, po.corr_offsets, po.corr_counts, po.corr_pos, po.corr_delta, po.status};
  for (size_t i = 0; i < sizeof dev / sizeof dev[0]; i++) CHECK_HIP(hipFree(dev[i]));
  free(bytes);
  free(offs);
  free(lens);
  free(expect);
  free(got);
  free(status);
  return bad_status || bad_chunks ? 1 : 0;
}

static int raycast_scene(r3d_hip_ctx *ctx) {
  size_t ns, n8, nf;
  uint8_t *sc = slurp("scene.bin", &ns);
  uint8_t *expect8 = slurp("expect_rgba8.bin", &n8);
  float *expectf = slurp("expect_rgba.bin", &nf);
  if (!sc || !expect8 || !expectf || ns < 24u) return 40;
  uint32_t hdr[6];
  memcpy(hdr, sc, sizeof hdr);
  const uint32_t nx = hdr[0], ny = hdr[1], nz = hdr[2], w = hdr[3], h = hdr[4], nviews = hdr[5];
  const size_t nvox = (size_t)nx * ny * nz, npx = (size_t)w * h;
  if (ns != 24u + sizeof(r3d_frame_params) * nviews + 1024u + nvox || n8 != 4u * npx || nf != 16u * npx) return 41;
  r3d_frame_params *views = malloc(sizeof(r3d_frame_params) * nviews);
  uint8_t tf[256][4];
  if (!views) return 42;
  memcpy(views, sc + 24u, sizeof(r3d_frame_params) * nviews);
  memcpy(tf, sc + 24u + sizeof(r3d_frame_params) * nviews, sizeof tf);
  const uint8_t *voxels = sc + 24u + sizeof(r3d_frame_params) * nviews + 1024u;

  const r3d_volume_desc desc = {nx, ny, nz, 128u, 1.0f};
  r3d_hip_volume *vol = NULL;
  CHECK_R3D(r3d_hip_volume_create(ctx, &desc, 0u, &vol));
  CHECK_R3D(r3d_hip_volume_upload(ctx, vol, voxels, R3D_SRC_HOST, NULL));
  CHECK_R3D(r3d_hip_set_transfer(ctx, (const uint8_t(*)[4])tf));
  float *d_rgba;
  uint8_t *d_rgba8;
  CHECK_HIP(hipMalloc((void **)&d_rgba, 16u * npx));
  CHECK_HIP(hipMalloc((void **)&d_rgba8, 4u * npx));
  CHECK_HIP(hipMemset(d_rgba, 0, 16u * npx));
  CHECK_HIP(hipMemset(d_rgba8, 0, 4u * npx));
  const r3d_hip_image img = {d_rgba, d_rgba8, w, h};
  r3d_raycast_opts opts;
  memset(&opts, 0, sizeof opts);
  opts.alpha_cutoff = 0.99f;
  r3d_hip_stats st;
  CHECK_R3D(r3d_hip_raycast(ctx, vol, views, nviews, &img, &opts, &st, NULL));
  uint8_t *got8 = malloc(4u * npx);
  float *gotf = malloc(16u * npx);
  if (!got8 || !gotf) return 43;
  CHECK_HIP(hipMemcpy(got8, d_rgba8, 4u * npx, hipMemcpyDeviceToHost));
  CHECK_HIP(hipMemcpy(gotf, d_rgba, 16u * npx, hipMemcpyDeviceToHost));
  size_t bad8 = 0;
  double maxerr = 0.0;
  for (size_t i = 0; i < 4u * npx; i++) {
    bad8 += got8[i] != expect8[i];
    const double e = fabs((double)gotf[i] - (double)expectf[i]);
    maxerr = e > maxerr ? e : maxerr;
  }
  printf("raycast: %ux%u, %u views, %llu samples, %llu shaded, %.0f us on the GPU, %zu RGBA8 bytes differ, "
         "max abs %.3g\n",
         w, h, nviews, (unsigned long long)st.samples, (unsigned long long)st.shaded, st.gpu_raycast_ns / 1e3, bad8,
         maxerr);
  if (spit("got_rgba8.bin", got8, 4u * npx) || spit("got_rgba.bin", gotf, 16u * npx)) return 44;
  CHECK_HIP(hipFree(d_rgba));
  CHECK_HIP(hipFree(d_rgba8));
  r3d_hip_volume_destroy(vol);
  free(sc);
  free(expect8);
  free(expectf);
  free(views);
  free(got8);
  free(gotf);
  return bad8 || maxerr > 1e-5 || st.samples == 0 ? 1 : 0;
}

int main(int argc, char **argv) {
  if (argc != 2) {
    fprintf(stderr, "usage: %s <dir>\n", argv[0]);
    return 2;
  }
  g_dir = argv[1];
  r3d_hip_ctx *ctx = NULL;
  const int rc = r3d_hip_init(&ctx, 0);
  if (rc != R3D_OK) {
    fprintf(stderr, "r3d_hip_init: %d\n", rc);
    return 3;
  }
  const int d = decode_golden(ctx);
  const int r = raycast_scene(ctx);
  r3d_hip_destroy(ctx);
  printf("decode %d raycast %d\n", d, r);
  return d ? d : r;
}
