/* oracle_deblock.c -- TEST INFRASTRUCTURE ONLY (the oracle): a plain-C
 * restatement of 3ddct's decode-side deblocking post-filter for u8 data
 * (tools/3ddct/deblock.c): the boundary sample filter (deblock.c:26-46), the
 * thresholds (:48-67), the chunk API (a chunk and its six decoded face
 * neighbours, :69-130) and the volume API (every internal 16-aligned plane in
 * place, x then y then z seams, :132-203). FP32 with no contraction, roundf,
 * clamp to [0, 255].
 *
 * Pinned against the reference deblock.c compiled where it lies
 * (oracle/_ref/libdeblock_ref.so) through tests/golden/dct3d_deblock.npz, and
 * by the checks of the reference's tools/3ddct/test_deblock.c (seam energy
 * falls, error against the original does not grow, a genuine edge on a block
 * boundary survives).
 */
#include <math.h>
#include <string.h>

#include "r3d_oracle.h"

#define DN 16
#define DIDX(z, y, x) ((((size_t)(z)) * DN + (size_t)(y)) * DN + (size_t)(x))

static float sample_delta(float p1, float p0, float q0, float q1, float alpha, float beta, float *d1) {
  *d1 = 0.0f;
  const float d = q0 - p0, ad = fabsf(d);
  if (ad >= alpha || ad == 0.0f) return 0.0f;
  if (fabsf(p1 - p0) >= beta) return 0.0f;
  if (fabsf(q1 - q0) >= beta) return 0.0f;
  float d0 = 0.40f * d;
  const float w = 1.0f - ad / alpha;
  d0 *= w;
  *d1 = 0.4f * d0;
  return d0;
}

static void limits(float step, float range, float strength, float *alpha, float *beta) {
  if (step > 0.0f) {
    *alpha = 2.0f * step * strength;
    *beta = 1.2f * step * strength;
  } else {
    *alpha = 0.06f * range * strength;
    *beta = 0.04f * range * strength;
  }
}

static uint8_t to_u8(float v) { return (uint8_t)roundf(v < 0.0f ? 0.0f : v > 255.0f ? 255.0f : v); }

static void minmax(const uint8_t *c, float *mn, float *mx) {
  float lo = c[0], hi = lo;
  for (int i = 1; i < DN * DN * DN; i++) {
    const float v = c[i];
    if (v < lo) lo = v;
    if (v > hi) hi = v;
  }
  *mn = lo;
  *mx = hi;
}

/* One face of the chunk API: neighbour values pristine, centre values
 * running (out); only the centre's two layers are written. */
static void face(const uint8_t *nbr, uint8_t *out, int axis, int lo_face, float step, float strength, float cmn,
                 float cmx) {
  float range = 1.0f;
  if (!(step > 0.0f)) {
    float nmn, nmx;
    minmax(nbr, &nmn, &nmx);
    range = (cmx > nmx ? cmx : nmx) - (cmn < nmn ? cmn : nmn);
  }
  if (!(range > 0.0f)) return;
  float alpha, beta;
  limits(step, range, strength, &alpha, &beta);
  for (int a = 0; a < DN; a++)
    for (int b = 0; b < DN; b++) {
      /* depth 0 / 1 from the face, for centre (c) and neighbour (n) */
      const int c0d = lo_face ? 0 : DN - 1, c1d = lo_face ? 1 : DN - 2;
      const int n0d = lo_face ? DN - 1 : 0, n1d = lo_face ? DN - 2 : 1;
      size_t ci0, ci1, ni0, ni1;
      if (axis == 0) {
        ci0 = DIDX(a, b, c0d), ci1 = DIDX(a, b, c1d), ni0 = DIDX(a, b, n0d), ni1 = DIDX(a, b, n1d);
      } else if (axis == 1) {
        ci0 = DIDX(a, c0d, b), ci1 = DIDX(a, c1d, b), ni0 = DIDX(a, n0d, b), ni1 = DIDX(a, n1d, b);
      } else {
        ci0 = DIDX(c0d, a, b), ci1 = DIDX(c1d, a, b), ni0 = DIDX(n0d, a, b), ni1 = DIDX(n1d, a, b);
      }
      const float n0 = nbr[ni0], n1 = nbr[ni1], c0 = out[ci0], c1 = out[ci1];
      float p1, p0, q0, q1;
      if (lo_face) {
        p1 = n1, p0 = n0, q0 = c0, q1 = c1;
      } else {
        p1 = c1, p0 = c0, q0 = n0, q1 = n1;
      }
      float d1;
      const float d0 = sample_delta(p1, p0, q0, q1, alpha, beta, &d1);
      if (d0 == 0.0f) continue;
      const float v0 = lo_face ? c0 - d0 : c0 + d0, v1 = lo_face ? c1 - d1 : c1 + d1;
      out[ci0] = to_u8(v0);
      out[ci1] = to_u8(v1);
    }
}

/* dct3d_deblock_chunk_u8: nbr[6] = xlo, xhi, ylo, yhi, zlo, zhi (NULL: no
 * neighbour), applied in that order. */
void r3do_deblock_chunk_u8(const uint8_t *chunk, const uint8_t *const nbr[6], float step, float strength,
                           uint8_t *out) {
  memcpy(out, chunk, DN * DN * DN);
  if (!(strength > 0.0f)) return;
  float cmn = 0.0f, cmx = 0.0f;
  if (!(step > 0.0f)) minmax(chunk, &cmn, &cmx);
  for (int f = 0; f < 6; f++)
    if (nbr[f]) face(nbr[f], out, f / 2, (f & 1) == 0, step, strength, cmn, cmx);
}

static void filter4(uint8_t *v, size_t s, float alpha, float beta) { /* v at the seam's q0, stride s */
  const float p1 = v[-2 * (ptrdiff_t)s], p0 = v[-(ptrdiff_t)s], q0 = v[0], q1 = v[s];
  float d1;
  const float d0 = sample_delta(p1, p0, q0, q1, alpha, beta, &d1);
  if (d0 == 0.0f) return;
  v[-(ptrdiff_t)s] = to_u8(p0 + d0);
  v[0] = to_u8(q0 - d0);
  v[-2 * (ptrdiff_t)s] = to_u8(p1 + d1);
  v[s] = to_u8(q1 - d1);
}

/* dct3d_deblock_u8: dims z, y, x. */
void r3do_deblock_u8(uint8_t *vol, uint64_t nz, uint64_t ny, uint64_t nx, float step, float strength) {
  if (nz == 0 || ny == 0 || nx == 0 || !(strength > 0.0f)) return;
  float range = 1.0f;
  if (!(step > 0.0f)) { /* the data range over a coarse subsample (deblock.c:144-152) */
    float vmin = vol[0], vmax = vol[0];
    const uint64_t total = nz * ny * nx;
    uint64_t stride = total / 4096;
    if (stride == 0) stride = 1;
    for (uint64_t i = 0; i < total; i += stride) {
      const float v = vol[i];
      if (v < vmin) vmin = v;
      if (v > vmax) vmax = v;
    }
    range = vmax - vmin;
    if (!(range > 0.0f)) return;
  }
  float alpha, beta;
  limits(step, range, strength, &alpha, &beta);
  for (uint64_t z = 0; z < nz; z++)
    for (uint64_t y = 0; y < ny; y++)
      for (uint64_t x = DN; x + 1 < nx; x += DN) filter4(vol + (z * ny + y) * nx + x, 1, alpha, beta);
  for (uint64_t z = 0; z < nz; z++)
    for (uint64_t y = DN; y + 1 < ny; y += DN)
      for (uint64_t x = 0; x < nx; x++) filter4(vol + (z * ny + y) * nx + x, nx, alpha, beta);
  for (uint64_t z = DN; z + 1 < nz; z += DN)
    for (uint64_t y = 0; y < ny; y++)
      for (uint64_t x = 0; x < nx; x++) filter4(vol + (z * ny + y) * nx + x, ny * nx, alpha, beta);
}
