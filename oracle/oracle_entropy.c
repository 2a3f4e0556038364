/* oracle_entropy.c -- TEST INFRASTRUCTURE ONLY (the oracle): a plain-C
 * restatement of the 3ddct entropy stage, the part of decode_float that turns
 * a blob into its header, quantised levels and corrections, and the matching
 * part of the encoder (tools/3ddct/dct3d.c:156-405, 414-433, 499-593).
 *
 * Pinned against the reference's own parser and encoder through the golden
 * blobs (tests/golden/dct3d_golden.npz: 40 reference-encoded blobs with the
 * reference parser's header, levels and corrections) and, where the reference
 * builds (oracle/_ref), against ref_parse_u8 on corrupted blobs.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>
#include <time.h>

#include "r3d_oracle.h"

#define N3 4096
#define TOP (1u << 24)

/* ---- binary range coder, decode side (dct3d.c:209-239) ---- */
typedef struct {
  const uint8_t *p;
  size_t n, at;
  uint32_t code, range;
} rdec;

static uint8_t next_byte(rdec *d) { return d->at < d->n ? d->p[d->at++] : 0; }

static void rdec_init(rdec *d, const uint8_t *p, size_t n) {
  d->p = p;
  d->n = n;
  d->at = 0;
  d->code = 0;
  d->range = 0xFFFFFFFFu;
  for (int i = 0; i < 5; i++) d->code = (d->code << 8) | next_byte(d);
}

static void rdec_norm(rdec *d) {
  while (d->range < TOP) {
    d->code = (d->code << 8) | next_byte(d);
    d->range <<= 8;
  }
}

/* adaptive bit: p0 = P(bit == 0) in 1/4096, adapting by 1/16 of the gap */
static int rdec_bit(rdec *d, uint16_t *p0) {
  const uint32_t r0 = (d->range >> 12) * *p0;
  int bit;
  if (d->code < r0) {
    d->range = r0;
    *p0 = (uint16_t)(*p0 + ((4096u - *p0) >> 4));
    bit = 0;
  } else {
    d->code -= r0;
    d->range -= r0;
    *p0 = (uint16_t)(*p0 - (*p0 >> 4));
    bit = 1;
  }
  rdec_norm(d);
  return bit;
}

static int rdec_bypass(rdec *d) {
  d->range >>= 1;
  const int bit = d->code >= d->range;
  if (bit) d->code -= d->range;
  rdec_norm(d);
  return bit;
}

static uint32_t rdec_bits(rdec *d, int k) { /* k raw bits, most significant first */
  uint32_t v = 0;
  while (k > 16) {
    v = (v << 16) | rdec_bits(d, 16);
    k -= 16;
  }
  if (k == 0) return v;
  d->range >>= k;
  uint32_t q = d->code / d->range;
  const uint32_t m = (1u << k) - 1u;
  if (q > m) q = m;
  d->code -= q * d->range;
  rdec_norm(d);
  return (v << k) | q;
}

static uint32_t rdec_eg(rdec *d) { /* order-0 Exp-Golomb, dct3d.c:247-251 */
  uint32_t nb = 0;
  while (rdec_bypass(d))
    if (++nb > 31) return 0;
  return nb ? ((1u << nb) | rdec_bits(d, (int)nb)) - 1u : 0u;
}

static long rdec_seg(rdec *d) { /* zig-zag signed, dct3d.c:258-261 */
  const uint32_t z = rdec_eg(d);
  return (z & 1u) ? -(long)((z + 1u) >> 1) : (long)(z >> 1);
}

static float rdec_fval(rdec *d) { /* dct3d.c:272-275 */
  if (!rdec_bypass(d)) return (float)rdec_seg(d);
  const uint32_t b = rdec_bits(d, 32);
  float f;
  memcpy(&f, &b, 4);
  return f;
}

/* ---- coefficient contexts (dct3d.c:281-319) ---- */
/* Trained priors P(bit == 0) in 1/4096 (dct3d.c:290-299); 0 would mean 2048. */
static const uint16_t PRIOR_SIG[32] = {1980, 575,  846,  743,  3916, 3707, 3515, 1463, 3969, 3753, 3527,
                                       2313, 4002, 3826, 3635, 2819, 4039, 3851, 3613, 2974, 3985, 3840,
                                       3703, 3475, 4021, 3991, 3951, 3969, 4065, 4024, 2048, 2048};
static const uint16_t PRIOR_MAG[12] = {1893, 1217, 908, 728, 610, 527, 465, 416, 377, 346, 320, 297};
static const uint16_t PRIOR_EOB[14] = {74, 1, 1, 1, 1, 5, 109, 504, 772, 1063, 1458, 2326, 4095, 2048};

typedef struct {
  uint16_t sig[32], mag[12], eob[14];
} cctx;

static void cctx_init(cctx *c) {
  for (int i = 0; i < 32; i++) c->sig[i] = PRIOR_SIG[i] ? PRIOR_SIG[i] : 2048;
  for (int i = 0; i < 12; i++) c->mag[i] = PRIOR_MAG[i] ? PRIOR_MAG[i] : 2048;
  for (int i = 0; i < 14; i++) c->eob[i] = PRIOR_EOB[i] ? PRIOR_EOB[i] : 2048;
}

/* ascending-L1 scan, ties by index (dct3d.c:336-351), and the 8 bands of the
 * L1 frequency (dct3d.c:352-355) */
static uint16_t g_scan[N3];
static uint8_t g_band[N3];
static int g_tables_ready;
static void tables_init(void) {
  if (g_tables_ready) return;
  int start[47] = {0};
  for (int i = 0; i < N3; i++) start[(i >> 8) + ((i >> 4) & 15) + (i & 15) + 1]++;
  for (int f = 1; f < 47; f++) start[f] += start[f - 1];
  for (int i = 0; i < N3; i++) {
    const int f = (i >> 8) + ((i >> 4) & 15) + (i & 15);
    g_scan[start[f]++] = (uint16_t)i;
    const int b = f * 8 / 48;
    g_band[i] = (uint8_t)(b > 7 ? 7 : b);
  }
  g_tables_ready = 1;
}

static uint32_t rdec_eob(rdec *d, cctx *c) { /* dct3d.c:365-371, kmax = 13 */
  int k = 0;
  while (k < 13 && rdec_bit(d, &c->eob[k])) k++;
  if (k <= 1) return (uint32_t)k;
  return (1u << (k - 1)) | rdec_bits(d, k - 1);
}

static uint32_t rdec_mag(rdec *d, cctx *c) { /* unary ladder + EG tail, dct3d.c:322-329 */
  uint32_t k = 0;
  for (; k < 11; k++)
    if (!rdec_bit(d, &c->mag[k])) return k + 1u;
  if (!rdec_bit(d, &c->mag[11])) return 12u;
  uint32_t nb = 0;
  while (rdec_bypass(d))
    if (++nb > 31) break;
  /* nb can reach 32 on a corrupt stream: the shift count is taken mod 32, as
   * the x86 shift instruction the reference compiles to does (dct3d.c:326) */
  const uint32_t x = nb ? ((1u << (nb & 31u)) | rdec_bits(d, (int)nb)) - 1u : 0u;
  return 11u + x + 1u;
}

/* 1: parsed; 0: malformed (the reference's decode_float returns 0); -1: a
 * level beyond int16 (parsed, levels clamped: unsupported by the int16 level
 * formats). */
int r3do_parse_blob(const uint8_t *blob, size_t len, uint32_t want_dtype, r3do_blob_hdr *h, int16_t lvl[4096],
                    uint32_t *ncorr_out, uint16_t *corr_pos, float *corr_delta) {
  *ncorr_out = 0;
  if (len < 2 || blob[0] != 0xD3u || blob[1] != want_dtype) return 0;
  tables_init();
  rdec d;
  rdec_init(&d, blob + 2, len - 2);
  h->vmin = rdec_fval(&d);
  h->vspan = rdec_fval(&d);
  h->qstep = rdec_fval(&d);
  h->dcq = (int32_t)rdec_seg(&d);
  const float dc = (float)h->dcq / 16.0f;
  if (!(h->vspan > 0.0f) || !(h->qstep > 0.0f) || !isfinite(h->vmin) || !isfinite(dc)) return 0;
  uint16_t flag = 2048;
  const int has_corr = rdec_bit(&d, &flag);
  /* levels (dct3d.c:390-405): EOB, then per scan position a significance bit
   * in context (band, min(3, popcount of the last 16 decisions)), magnitude, sign */
  cctx c;
  cctx_init(&c);
  memset(lvl, 0, sizeof(int16_t) * N3);
  int wide = 0; /* a magnitude beyond int16 (the reference keeps levels as floats) */
  uint32_t eob = rdec_eob(&d, &c);
  if (eob > N3) eob = N3;
  uint32_t recent = 0;
  for (uint32_t p = 0; p < eob; p++) {
    const uint32_t idx = g_scan[p];
    int dens = __builtin_popcount(recent & 0xFFFFu);
    if (dens > 3) dens = 3;
    const int sig = p + 1 == eob ? 1 : rdec_bit(&d, &c.sig[g_band[idx] * 4 + dens]);
    recent = (recent << 1) | (uint32_t)sig;
    if (!sig) continue;
    const uint32_t m = rdec_mag(&d, &c);
    const int neg = rdec_bypass(&d);
    const uint32_t lim = neg ? 32768u : 32767u;
    if (m > lim) wide = 1;
    lvl[idx] = (int16_t)(neg ? -(int32_t)(m > lim ? lim : m) : (int32_t)(m > lim ? lim : m));
  }
  if (has_corr) { /* dct3d.c:572-587 */
    uint32_t n = rdec_eg(&d) + 1u, pos = 0, nc = 0;
    if (n > N3) n = N3;
    float cq = rdec_fval(&d);
    if (!(cq > 0.0f) || !isfinite(cq)) cq = 1.0f;
    for (uint32_t i = 0; i < n; i++) {
      pos += rdec_eg(&d);
      const int neg = rdec_bypass(&d);
      const uint32_t mag = rdec_eg(&d) + 1u;
      if (pos >= N3) break;
      const float delta = (float)mag * cq;
      corr_pos[nc] = (uint16_t)pos;
      corr_delta[nc] = neg ? -delta : delta;
      nc++;
    }
    *ncorr_out = nc;
  }
  return wide ? -1 : 1;
}

/* ---- encode side (dct3d.c:180-243, 357-388, 499-537), no corrections ---- */
typedef struct {
  uint8_t *out;
  size_t cap, n;
  uint64_t low;
  uint32_t range;
  uint8_t cache;
  uint64_t pending;
} renc;

static void renc_put(renc *e, uint8_t b) {
  if (e->n < e->cap) e->out[e->n] = b;
  e->n++;
}

static void renc_shift(renc *e) { /* carry propagation through pending 0xFF bytes */
  if ((uint32_t)(e->low >> 32) != 0 || e->low < 0xFF000000ull) {
    const uint8_t carry = (uint8_t)(e->low >> 32);
    do {
      renc_put(e, (uint8_t)(e->cache + carry));
      e->cache = 0xFF;
    } while (--e->pending);
    e->cache = (uint8_t)(e->low >> 24);
  }
  e->pending++;
  e->low = (e->low << 8) & 0xFFFFFFFFull;
}

static void renc_norm(renc *e) {
  while (e->range < TOP) {
    renc_shift(e);
    e->range <<= 8;
  }
}

static void renc_bit(renc *e, uint16_t *p0, int bit) {
  const uint32_t r0 = (e->range >> 12) * *p0;
  if (!bit) {
    e->range = r0;
    *p0 = (uint16_t)(*p0 + ((4096u - *p0) >> 4));
  } else {
    e->low += r0;
    e->range -= r0;
    *p0 = (uint16_t)(*p0 - (*p0 >> 4));
  }
  renc_norm(e);
}

static void renc_bypass(renc *e, int bit) {
  e->range >>= 1;
  if (bit) e->low += e->range;
  renc_norm(e);
}

static void renc_bits(renc *e, uint32_t v, int k) {
  while (k > 16) {
    renc_bits(e, (v >> (k - 16)) & 0xFFFFu, 16);
    k -= 16;
  }
  if (k == 0) return;
  e->range >>= k;
  e->low += (uint64_t)(v & ((1u << k) - 1u)) * e->range;
  renc_norm(e);
}

static void renc_eg(renc *e, uint32_t v) {
  uint32_t nb = 0, t = v + 1u;
  while (t > 1) {
    t >>= 1;
    nb++;
  }
  for (uint32_t i = 0; i < nb; i++) renc_bypass(e, 1);
  renc_bypass(e, 0);
  if (nb) renc_bits(e, (v + 1u) & ((1u << nb) - 1u), (int)nb);
}

static void renc_seg(renc *e, long v) { renc_eg(e, v < 0 ? (uint32_t)(-2 * v - 1) : (uint32_t)(2 * v)); }

static void renc_fval(renc *e, float f) {
  const float r = truncf(f);
  if (r == f && fabsf(f) < 16777216.0f) {
    renc_bypass(e, 0);
    renc_seg(e, (long)r);
  } else {
    renc_bypass(e, 1);
    uint32_t b;
    memcpy(&b, &f, 4);
    renc_bits(e, b, 32);
  }
}

size_t r3do_encode_blob(const r3do_blob_hdr *h, const int16_t lvl[4096], uint32_t dtype, uint8_t *out, size_t cap) {
  tables_init();
  if (cap < 2) return 0;
  out[0] = 0xD3u;
  out[1] = (uint8_t)dtype;
  renc e = {out + 2, cap - 2, 0, 0, 0xFFFFFFFFu, 0, 1};
  renc_fval(&e, h->vmin);
  renc_fval(&e, h->vspan);
  renc_fval(&e, h->qstep);
  renc_seg(&e, (long)h->dcq);
  uint16_t flag = 2048;
  renc_bit(&e, &flag, 0);
  cctx c;
  cctx_init(&c);
  uint32_t eob = 0;
  for (uint32_t p = N3; p-- > 0;)
    if (lvl[g_scan[p]] != 0) {
      eob = p + 1;
      break;
    }
  /* EOB: k = bit length of eob; k context bits 1, a 0 unless k == 13, then the
   * low k - 1 bits raw (dct3d.c:357-363) */
  int k = 0;
  while ((1u << k) <= eob) k++;
  for (int i = 0; i < k; i++) renc_bit(&e, &c.eob[i], 1);
  if (k < 13) renc_bit(&e, &c.eob[k], 0);
  if (k > 1) renc_bits(&e, eob & ((1u << (k - 1)) - 1u), k - 1);
  uint32_t recent = 0;
  for (uint32_t p = 0; p < eob; p++) {
    const uint32_t idx = g_scan[p];
    const int v = lvl[idx];
    int dens = __builtin_popcount(recent & 0xFFFFu);
    if (dens > 3) dens = 3;
    if (p + 1 != eob) renc_bit(&e, &c.sig[g_band[idx] * 4 + dens], v != 0);
    recent = (recent << 1) | (uint32_t)(v != 0);
    if (!v) continue;
    /* magnitude (dct3d.c:310-319): unary over 11 contexts, escape, EG tail */
    uint32_t m = (uint32_t)(v < 0 ? -v : v), rest = m - 1u, j = 0;
    while (j < 11 && rest > 0) {
      renc_bit(&e, &c.mag[j], 1);
      rest--;
      j++;
    }
    if (rest == 0) {
      renc_bit(&e, &c.mag[j], 0);
    } else {
      renc_bit(&e, &c.mag[11], 1);
      renc_eg(&e, rest);
    }
    renc_bypass(&e, v < 0);
  }
  for (int i = 0; i < 5; i++) renc_shift(&e);
  return e.n + 2;
}

/* Blob batch decode, the CPU shape of r3d_shard_decode_region
 * (src/shard/shardio.c:269-306 calling dct3d_decode_u8 per chunk): blob i is
 * bytes[offsets[i] .. + lengths[i]], parsed, dequantised with its own
 * qstep table and decoded into out + i*4096. status[i] as r3do_parse_blob;
 * a chunk with status != 1 is zero-filled. Runs on nthreads threads. */
void r3do_decode_blobs(uint32_t n, const uint8_t *bytes, const uint64_t *offsets, const uint32_t *lengths,
                       uint32_t dtype, uint8_t *out, int32_t *status, int nthreads) {
  if (nthreads < 1) nthreads = 1;
  tables_init();
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads)
  for (uint32_t i = 0; i < n; i++) {
    r3do_blob_hdr h;
    int16_t lvl[N3];
    uint16_t cp[N3];
    float cd[N3], step[46];
    uint32_t nc = 0;
    uint8_t *o = out + (size_t)i * N3;
    int st = r3do_parse_blob(bytes + offsets[i], lengths[i], dtype, &h, lvl, &nc, cp, cd);
    if (st == 1) {
      r3do_step_table_qstep(h.qstep, step);
      const r3do_chunk_hdr ch = {h.vmin, h.vspan, h.dcq, 0};
      if (!r3do_decode_chunk(lvl, &ch, step, cp, cd, nc, o, NULL)) st = 0;
    }
    if (st != 1) memset(o, 0, N3);
    status[i] = st;
  }
}

/* The timing loop of tools/3ddct/perf.c:55-60 for any decoder of
 * dct3d_decode_u8's signature (the reference's own, from the build in
 * oracle/_ref): `iters` passes over the n blobs, each pass split over
 * nthreads OpenMP threads. Returns the wall seconds; *ok counts successful
 * decodes and *sink sums the first voxel of each, so no decode is elided. */
double r3do_time_decoder(int (*decode)(const uint8_t *, size_t, uint8_t *), const uint8_t *bytes,
                         const uint64_t *offsets, const uint32_t *lengths, uint32_t n, uint32_t iters, int nthreads,
                         uint64_t *ok, uint64_t *sink) {
  struct timespec t0, t1;
  uint64_t good = 0, acc = 0;
  if (nthreads < 1) nthreads = 1;
  clock_gettime(CLOCK_MONOTONIC, &t0);
  for (uint32_t it = 0; it < iters; it++) {
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads) reduction(+ : good, acc)
    for (int64_t i = 0; i < (int64_t)n; i++) {
      uint8_t out[4096];
      good += (uint64_t)(decode(bytes + offsets[i], lengths[i], out) == 1);
      acc += out[0];
    }
  }
  clock_gettime(CLOCK_MONOTONIC, &t1);
  *ok = good;
  *sink = acc;
  return (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
}
