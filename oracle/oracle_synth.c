/* oracle_synth.c -- TEST INFRASTRUCTURE ONLY (the oracle; see r3d_oracle.h).
 *
 * Host restatement of the synthetic sheet-volume generator (the workloads'
 * input; the reference has no such generator, tools/mkvol/main.c:10-29 is the
 * pattern). Pinned sequence, shared with the HIP generator:
 *
 *   idx = (z*ny + y)*nx + x;  h = mix64(seed ^ (idx * 0xD1B54A32D192ED03))
 *   a = low 32 bits of h, b = high 32 bits, c = low 32 bits of mix64(h)
 *   bg = 20 + (c % 11) - 5
 *   dx = (x + 0.5) - nx/2, dy = (y + 0.5) - ny/2, r = sqrtf(dx*dx + dy*dy)
 *   rr = r - 3 * sin_turns(z / 40)
 *   k = (int)floorf((rr - 20) / 9 + 0.5); sheet if 0 <= k < nsheets and
 *   |rr - (20 + 9k)| < 1.5, with value 160 + off_k,
 *   off_k = (mix64(seed ^ (0x5EED0000 + k)) % 61) - 30
 *   u1 = ((a >> 8) + 0.5) * 2^-24, u2 = ((b >> 8) + 0.5) * 2^-24
 *   g = sqrtf(-2 * log2f(u1) * LN2) * cos_turns(u2)      (Box-Muller)
 *   voxel = clamp(roundf(base + 8 * g), 0, 255)
 * mix64 is the splitmix64 finaliser.
 */
#include <math.h>
#include <stdint.h>

#include "oracle_math.h"
#include "r3d_oracle.h"

static inline uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

void r3do_synth_sheets(uint8_t *vol, uint32_t nx, uint32_t ny, uint32_t nz, uint32_t nsheets,
                       uint64_t seed, uint32_t z0, uint32_t z1, int nthreads) {
  int32_t off[256];
  if (nsheets > 256) nsheets = 256;
  for (uint32_t k = 0; k < nsheets; k++) off[k] = (int32_t)(mix64(seed ^ (0x5EED0000ull + k)) % 61u) - 30;
  const float cx = (float)nx * 0.5f, cy = (float)ny * 0.5f;
  const float LN2 = 0x1.62e430p-1f;
  if (z1 > nz) z1 = nz;
  if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
  for (int64_t z = z0; z < (int64_t)z1; z++) {
    float ws, wc;
    om_sincos_turns((float)z / 40.0f, &ws, &wc);
    const float wave = 3.0f * ws;
    for (uint32_t y = 0; y < ny; y++) {
      const float dy = ((float)y + 0.5f) - cy;
      for (uint32_t x = 0; x < nx; x++) {
        uint64_t idx = ((uint64_t)z * ny + y) * nx + x;
        uint64_t h = mix64(seed ^ (idx * 0xD1B54A32D192ED03ull));
        uint32_t a = (uint32_t)h, b = (uint32_t)(h >> 32), c = (uint32_t)mix64(h);
        int32_t base = 20 + (int32_t)(c % 11u) - 5;
        const float dx = ((float)x + 0.5f) - cx;
        float r = sqrtf(dx * dx + dy * dy);
        float rr = r - wave;
        float kf = floorf((rr - 20.0f) / 9.0f + 0.5f);
        if (kf >= 0.0f && kf < (float)nsheets) {
          int32_t k = (int32_t)kf;
          if (fabsf(rr - (20.0f + 9.0f * (float)k)) < 1.5f) base = 160 + off[k];
        }
        float u1 = ((float)(a >> 8) + 0.5f) * 0x1p-24f;
        float u2 = ((float)(b >> 8) + 0.5f) * 0x1p-24f;
        float s2, c2;
        om_sincos_turns(u2, &s2, &c2);
        float g = sqrtf(-2.0f * (om_log2f(u1) * LN2)) * c2;
        float v = roundf((float)base + 8.0f * g);
        v = om_clamp(v, 0.0f, 255.0f);
        vol[idx] = (uint8_t)v;
      }
    }
  }
}
