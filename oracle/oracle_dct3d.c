/* oracle_dct3d.c -- TEST INFRASTRUCTURE ONLY (the oracle; see r3d_oracle.h).
 *
 * Restates the arithmetic of tools/3ddct/dct3d.c for the decode back half and
 * the forward transform + quantiser. Compiled with -ffp-contract=off and
 * without -ffast-math: every product and sum is rounded separately, in the
 * reference's order. That is the pinned sequence the HIP decoder must match
 * bit for bit.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "r3d_oracle.h"

#define N 16
#define N3 4096

static float g_cm[N][N];
static int g_cm_ready;

/* dct3d.c:38-49: cm[k][n] = (float)(ck * cos(M_PI * (2n+1) * k / (2N))). */
void r3do_dct_basis(float cm[256]) {
  for (int k = 0; k < N; ++k) {
    double ck = (k == 0) ? sqrt(1.0 / N) : sqrt(2.0 / N);
    for (int n = 0; n < N; ++n)
      cm[k * N + n] = (float)(ck * cos(3.14159265358979323846 * (2.0 * n + 1.0) * k / (2.0 * N)));
  }
}

static void basis_init(void) {
  if (__atomic_load_n(&g_cm_ready, __ATOMIC_ACQUIRE)) return;
  float cm[256];
  r3do_dct_basis(cm);
  memcpy(g_cm, cm, sizeof cm);
  __atomic_store_n(&g_cm_ready, 1, __ATOMIC_RELEASE);
}

/* dct3d.c:131-143 with qstep applied: step[f] = qstep * powf(1 + f, 0.65f). */
void r3do_step_table_qstep(float qstep, float step46[46]) {
  for (int f = 0; f <= 45; ++f) step46[f] = qstep * powf(1.0f + (float)f, 0.65f);
}

void r3do_step_table_linear(float step46[46]) {
  for (int f = 0; f <= 45; ++f) step46[f] = 1.0f + 2.0f * (float)f;
}

/* dct3d.c:149-154: r = ((|l| - 1) * step + 0.8 * step) + 0.4 * step. */
static inline float deq_one(float lv, float step) {
  if (lv == 0.0f) return 0.0f;
  float a = fabsf(lv);
  float r = (a - 1.0f) * step + 0.80f * step + 0.40f * step;
  return lv < 0.0f ? -r : r;
}

/* dct3d.c:144-148: dead-zone quantiser. */
static inline float quant_one(float c, float step) {
  float dz = 0.80f * step, a = fabsf(c);
  float lv = 0.0f;
  if (a >= dz) lv = floorf((a - dz) / step + 1.0f);
  return c < 0.0f ? -lv : lv;
}

/* dct3d.c:61-68: DCT-III of one line; even terms summed in ascending k into
 * e[n], odd into o[n], zero inputs skipped; out[n] = e+o, out[15-n] = e-o. */
static void idct_line(const float *in, float *out) {
  float e[8], o[8];
  for (int n = 0; n < 8; ++n) { e[n] = 0.0f; o[n] = 0.0f; }
  for (int k = 0; k < N; k += 2) {
    float v = in[k];
    if (v != 0.0f)
      for (int n = 0; n < 8; ++n) e[n] += g_cm[k][n] * v;
  }
  for (int k = 1; k < N; k += 2) {
    float v = in[k];
    if (v != 0.0f)
      for (int n = 0; n < 8; ++n) o[n] += g_cm[k][n] * v;
  }
  for (int n = 0; n < 8; ++n) { out[n] = e[n] + o[n]; out[N - 1 - n] = e[n] - o[n]; }
}

/* dct3d.c:52-59: DCT-II of one line via s = in[n] + in[15-n], d = in[n] - in[15-n]. */
static void fdct_line(const float *in, float *out) {
  float s[8], d[8];
  for (int n = 0; n < 8; ++n) { s[n] = in[n] + in[N - 1 - n]; d[n] = in[n] - in[N - 1 - n]; }
  for (int k = 0; k < N; k += 2) {
    float a = 0.0f;
    for (int n = 0; n < 8; ++n) a += g_cm[k][n] * s[n];
    out[k] = a;
  }
  for (int k = 1; k < N; k += 2) {
    float a = 0.0f;
    for (int n = 0; n < 8; ++n) a += g_cm[k][n] * d[n];
    out[k] = a;
  }
}

/* dct3d.c:70-89: transform the 256 contiguous lines; all-zero lines -> 0. */
static void lines(const float *src, float *dst, int inverse) {
  for (int off = 0; off < N * N; ++off) {
    const float *v = src + off * N;
    float *o = dst + off * N;
    int nz = 0;
    for (int i = 0; i < N; ++i)
      if (v[i] != 0.0f) { nz = 1; break; }
    if (!nz) {
      for (int i = 0; i < N; ++i) o[i] = 0.0f;
      continue;
    }
    if (inverse) idct_line(v, o); else fdct_line(v, o);
  }
}

/* dct3d.c:93-104: axis rotation (z,y,x) -> (x,z,y): dst[(x*16+z)*16+y] = src[(z*16+y)*16+x]. */
static void rot(const float *src, float *dst) {
  for (int z = 0; z < N; ++z)
    for (int y = 0; y < N; ++y)
      for (int x = 0; x < N; ++x) dst[(x * N + z) * N + y] = src[(z * N + y) * N + x];
}

/* dct3d.c:113-118: three (line pass, rotate) rounds. */
void r3do_idct3(const float coef[4096], float blk[4096]) {
  basis_init();
  float a[N3], b[N3];
  lines(coef, a, 1); rot(a, b);
  lines(b, a, 1);    rot(a, b);
  lines(b, a, 1);    rot(a, blk);
}

/* dct3d.c:106-111. */
void r3do_fdct3(const float blk[4096], float coef[4096]) {
  basis_init();
  float a[N3], b[N3];
  lines(blk, a, 0); rot(a, b);
  lines(b, a, 0);   rot(a, b);
  lines(b, a, 0);   rot(a, coef);
}

/* dct3d.c:564-566 with step_tab[i] = step46[cz+cy+cx] (dct3d.c:138-143). */
void r3do_dequant(const int16_t lvl[4096], const float step46[46], float coef[4096]) {
  for (int cz = 0; cz < N; ++cz)
    for (int cy = 0; cy < N; ++cy)
      for (int cx = 0; cx < N; ++cx) {
        int i = (cz * N + cy) * N + cx;
        coef[i] = deq_one((float)lvl[i], step46[cz + cy + cx]);
      }
}

/* dct3d.c:439-442: clamp to [lo,hi], NaN -> lo. */
static inline float clampf_nan_lo(float v, float lo, float hi) {
  if (!(v >= lo)) return lo;
  return v > hi ? hi : v;
}

/* decode_float (dct3d.c:542-593) from the levels onward, then DEC_BODY
 * (dct3d.c:625-635) with roundf and the [0,255] clamp. */
int r3do_decode_chunk(const int16_t lvl[4096], const r3do_chunk_hdr *h, const float step46[46],
                      const uint16_t *corr_pos, const float *corr_delta, uint32_t ncorr,
                      uint8_t out[4096], float *pre_round) {
  float vmin = h->vmin, vspan = h->vspan;
  float dc = (float)h->dcq / 16.0f;
  if (!(vspan > 0.0f) || !isfinite(vmin) || !isfinite(dc)) return 0;
  float coef[N3], rb[N3], norm[N3];
  r3do_dequant(lvl, step46, coef);
  r3do_idct3(coef, rb);
  for (int i = 0; i < N3; ++i) norm[i] = rb[i] + dc;
  if (ncorr > N3) ncorr = N3;
  for (uint32_t c = 0; c < ncorr; ++c) {
    uint32_t pos = corr_pos[c];
    if (pos >= N3) break;
    norm[pos] += corr_delta[c];
  }
  const float inv = vspan / 255.0f;
  for (int i = 0; i < N3; ++i) {
    float v = norm[i] * inv + vmin;
    if (pre_round) pre_round[i] = v;
    float r = clampf_nan_lo(roundf(v), 0.0f, 255.0f);
    out[i] = (uint8_t)r;
  }
  return 1;
}

void r3do_decode_batch(uint32_t n, const int16_t *levels, const r3do_chunk_hdr *hdr,
                       const float *tables, uint32_t ntables, const uint32_t *corr_off,
                       const uint16_t *corr_pos, const float *corr_delta, uint8_t *out,
                       int32_t *status, int nthreads) {
  basis_init();
  if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads)
  for (int64_t i = 0; i < (int64_t)n; ++i) {
    uint32_t t = hdr[i].table;
    int ok = 0;
    if (t < ntables) {
      uint32_t c0 = corr_off ? corr_off[i] : 0, c1 = corr_off ? corr_off[i + 1] : 0;
      ok = r3do_decode_chunk(levels + (size_t)i * N3, &hdr[i], tables + (size_t)t * 46,
                             corr_off ? corr_pos + c0 : NULL, corr_off ? corr_delta + c0 : NULL,
                             c1 > c0 ? c1 - c0 : 0, out + (size_t)i * N3, NULL);
    }
    if (!ok) memset(out + (size_t)i * N3, 0, N3);
    if (status) status[i] = ok;
  }
}

/* ENC_BODY (dct3d.c:598-621) for u8 and encode_float (dct3d.c:431-472) up
 * to the quantiser. */
void r3do_encode_chunk(const uint8_t in[4096], const float step46[46], int16_t lvl[4096],
                       r3do_chunk_hdr *h) {
  basis_init();
  float vmin = (float)in[0], vmax = (float)in[0];
  for (int i = 1; i < N3; ++i) {
    float v = (float)in[i];
    if (v < vmin) vmin = v;
    if (v > vmax) vmax = v;
  }
  float vspan = vmax - vmin;
  if (!(vspan > 0.0f)) vspan = 1.0f;
  const float scale = 255.0f / vspan;
  float norm[N3], blk[N3], coef[N3];
  float sum = 0.0f;
  for (int i = 0; i < N3; ++i) { float w = ((float)in[i] - vmin) * scale; norm[i] = w; sum += w; }
  float dc = sum / (float)N3;
  long dcq_i = lroundf(dc * 16.0f);
  float dcq = (float)dcq_i / 16.0f;
  for (int i = 0; i < N3; ++i) blk[i] = norm[i] - dcq;
  r3do_fdct3(blk, coef);
  for (int cz = 0; cz < N; ++cz)
    for (int cy = 0; cy < N; ++cy)
      for (int cx = 0; cx < N; ++cx) {
        int i = (cz * N + cy) * N + cx;
        float l = quant_one(coef[i], step46[cz + cy + cx]);
        if (l > 32767.0f) l = 32767.0f;
        if (l < -32768.0f) l = -32768.0f;
        lvl[i] = (int16_t)l;
      }
  h->vmin = vmin;
  h->vspan = vspan;
  h->dcq = (int32_t)dcq_i;
  h->table = 0;
}

void r3do_encode_volume(const uint8_t *vol, uint32_t nx, uint32_t ny, uint32_t nz,
                        const float step46[46], int16_t *levels, r3do_chunk_hdr *hdr,
                        int nthreads) {
  basis_init();
  uint32_t cx = nx / 16, cy = ny / 16, cz = nz / 16;
  int64_t nch = (int64_t)cx * cy * cz;
  if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads)
  for (int64_t c = 0; c < nch; ++c) {
    uint32_t bx = (uint32_t)(c % cx), by = (uint32_t)((c / cx) % cy), bz = (uint32_t)(c / ((int64_t)cx * cy));
    uint8_t blk[N3];
    for (int z = 0; z < N; ++z)
      for (int y = 0; y < N; ++y)
        memcpy(blk + (z * N + y) * N,
               vol + (((size_t)(bz * 16 + z) * ny + (by * 16 + y)) * nx + bx * 16), N);
    r3do_encode_chunk(blk, step46, levels + (size_t)c * N3, &hdr[c]);
  }
}
