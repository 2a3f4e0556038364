// r3d_ctx.hip -- context lifetime, errors, transfer tables.
#include <math.h>
#include <new>
#include <stdlib.h>
#include <string.h>

#include "r3d_internal.h"

int r3d_fail(r3d_hip_ctx *ctx, int code, const char *fmt, ...) {
  if (ctx) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(ctx->err, sizeof ctx->err, fmt, ap);
    va_end(ap);
  }
  return code;
}

// cm[k][n] = (float)(ck * cos(pi*(2n+1)*k/32)) evaluated in double in the
// reference's operand order (tools/3ddct/dct3d.c:38-49).
static void build_basis(float cm[256]) {
  for (int k = 0; k < 16; ++k) {
    double ck = (k == 0) ? sqrt(1.0 / 16) : sqrt(2.0 / 16);
    for (int n = 0; n < 16; ++n)
      cm[k * 16 + n] = (float)(ck * cos(3.14159265358979323846 * (2.0 * n + 1.0) * k / (2.0 * 16)));
  }
}

extern "C" int r3d_hip_init(r3d_hip_ctx **out, int device) {
  if (!out) return R3D_ERR_INVALID;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return R3D_ERR_DEVICE;
  if (device < 0 || device >= ndev) return R3D_ERR_INVALID;
  r3d_hip_ctx *c = new (std::nothrow) r3d_hip_ctx();
  if (!c) return R3D_ERR_NOMEM;
  c->device = device;
  if (hipSetDevice(device) != hipSuccess) {
    delete c;
    return R3D_ERR_DEVICE;
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->num_cus = prop.multiProcessorCount;
  build_basis(c->basis);
  int rc = R3D_OK;
  if (hipMalloc(&c->d_tf, 256 * 4 * sizeof(float)) != hipSuccess ||
      hipMalloc(&c->d_counters, 8 * sizeof(unsigned long long)) != hipSuccess ||
      hipMalloc(&c->d_queue, 64 * sizeof(uint32_t)) != hipSuccess || hipMalloc(&c->d_dblim, 16) != hipSuccess) {
    rc = R3D_ERR_NOMEM;
  }
  if (rc == R3D_OK) rc = r3d::upload_basis(c);
  if (rc == R3D_OK) rc = r3d::upload_blob_tables(c);
  if (rc == R3D_OK) rc = r3d::upload_crc_tables(c);
  if (rc == R3D_OK) rc = r3d::parse_noalloc_setup(c);
  if (rc == R3D_OK && (hipEventCreate(&c->ev_ray[0]) != hipSuccess || hipEventCreate(&c->ev_ray[1]) != hipSuccess))
    rc = R3D_ERR_DEVICE;
  if (rc == R3D_OK) {
    // default transfer function: the grayscale identity ramp (r3d_tf_preset 0,
    // src/core/transfer.c:28-32), as the reference uploads before any TF is set
    float ramp[256 * 4];
    for (int i = 0; i < 256; i++)
      for (int ch = 0; ch < 4; ch++) ramp[i * 4 + ch] = (float)i * (1.0f / 255.0f);
    if (hipMemcpy(c->d_tf, ramp, sizeof ramp, hipMemcpyHostToDevice) != hipSuccess) rc = R3D_ERR_DEVICE;
  }
  if (rc != R3D_OK) {
    r3d_hip_destroy(c);
    return rc;
  }
  *out = c;
  return R3D_OK;
}

extern "C" void r3d_hip_destroy(r3d_hip_ctx *c) {
  if (!c) return;
  hipSetDevice(c->device);
  for (auto &s : c->slots) {
    if (s.ev) {
      hipEventSynchronize(s.ev);
      hipEventDestroy(s.ev);
    }
    if (s.host) hipHostFree(s.host);
    if (s.dev) hipFree(s.dev);
  }
  for (hipEvent_t e : c->ev_ray)
    if (e) hipEventDestroy(e);
  if (c->d_tf) hipFree(c->d_tf);
  if (c->d_counters) hipFree(c->d_counters);
  if (c->d_queue) hipFree(c->d_queue);
  if (c->d_partials) hipFree(c->d_partials);
  if (c->d_order) hipFree(c->d_order);
  if (c->d_gtf) hipFree(c->d_gtf);
  if (c->d_ws) hipFree(c->d_ws);
  if (c->d_dblim) hipFree(c->d_dblim);
  if (c->d_sort) hipFree(c->d_sort);
  for (uint32_t i = 0; i < c->nretired; i++) hipFree(c->retired[i]);
  delete c;
}

extern "C" const char *r3d_hip_last_error(const r3d_hip_ctx *ctx) { return ctx ? ctx->err : "no context"; }

extern "C" int r3d_hip_sync(r3d_hip_ctx *ctx, void *stream) {
  if (!ctx) return R3D_ERR_INVALID;
  R3D_HIP_CHECK(ctx, hipStreamSynchronize((hipStream_t)stream));
  return R3D_OK;
}

extern "C" int r3d_hip_dct_basis(const r3d_hip_ctx *ctx, float out[256]) {
  if (!ctx || !out) return R3D_ERR_INVALID;
  memcpy(out, ctx->basis, sizeof ctx->basis);
  return R3D_OK;
}

// RGBA8 unorm -> float (v * (1/255), the K255 constant the samplers use).
extern "C" int r3d_hip_set_transfer(r3d_hip_ctx *ctx, const uint8_t rgba[256][4]) {
  if (!ctx || !rgba) return R3D_ERR_INVALID;
  float t[256 * 4];
  for (int i = 0; i < 256; i++)
    for (int ch = 0; ch < 4; ch++) t[i * 4 + ch] = (float)rgba[i][ch] * (1.0f / 255.0f);
  R3D_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  R3D_HIP_CHECK(ctx, hipMemcpy(ctx->d_tf, t, sizeof t, hipMemcpyHostToDevice));
  r3d_scene_changed(ctx);
  return R3D_OK;
}

extern "C" int r3d_hip_set_transfer_f32(r3d_hip_ctx *ctx, const float *rgba) {
  if (!ctx || !rgba) return R3D_ERR_INVALID;
  for (int i = 0; i < 256 * 4; i++)
    if (!isfinite(rgba[i])) return r3d_fail(ctx, R3D_ERR_INVALID, "transfer entry %d is not finite", i);
  R3D_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  R3D_HIP_CHECK(ctx, hipMemcpy(ctx->d_tf, rgba, 256 * 4 * sizeof(float), hipMemcpyHostToDevice));
  r3d_scene_changed(ctx);
  return R3D_OK;
}

// For writes the library cannot see (a caller's LUT passed as opts->tf, the
// bricks tables written directly): every pane-cached view redraws.
extern "C" void r3d_hip_invalidate_panes(r3d_hip_ctx *ctx) {
  if (ctx) r3d_scene_changed(ctx);
}
