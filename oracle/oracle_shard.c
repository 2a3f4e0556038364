/* oracle_shard.c -- TEST INFRASTRUCTURE ONLY (the oracle): a plain-C
 * restatement of the zarr v3 sharding_indexed container of 3ddct chunks as
 * src/shard/shardio.c reads it (the format of tools/3ddct/tools/export/shard.h:
 * blobs back to back, then one (u64le offset, u64le nbytes) pair per 16^3
 * chunk in z-major order, then the crc32c of that index; (~0, ~0) marks a
 * missing chunk).
 *
 * Pinned against the reference reader and writer compiled where they lie
 * (oracle/_ref/libshard_ref.so): tests/golden/dct3d_shards.npz holds shards
 * the reference writer produced, region decodes by the reference
 * r3d_shard_decode_region (return code and bytes), and damaged shards
 * (bad crc, an overflowing index entry, corrupt payload, truncation) with the
 * reference's results. The CRC is also pinned to the published CRC-32C check
 * value (crc32c("123456789") = 0xE3069283).
 */
#include <string.h>

#include "r3d_oracle.h"

#define SH_GRID 64u
#define SH_CHUNK 16u
#define SH_N (SH_GRID * SH_GRID * SH_GRID)
#define SH_INDEX_BYTES ((uint64_t)SH_N * 16u + 4u)

static uint64_t le64(const uint8_t *p) {
  uint64_t v = 0;
  for (int i = 7; i >= 0; i--) v = (v << 8) | p[i];
  return v;
}

/* CRC-32C: reflected Castagnoli polynomial 0x82F63B78, initial value and
 * final xor 0xFFFFFFFF (shardio.c:39-52), one bit at a time. */
uint32_t r3do_crc32c(const uint8_t *p, size_t n) {
  uint32_t c = 0xFFFFFFFFu;
  while (n--) {
    c ^= *p++;
    for (int k = 0; k < 8; k++) c = (c >> 1) ^ (0x82F63B78u & (0u - (c & 1u)));
  }
  return ~c;
}

/* shardio.c:114-141: a shard shorter than its index, or whose index crc does
 * not match, is corrupt. 1 valid, 0 corrupt. */
int r3do_shard_valid(const uint8_t *shard, uint64_t n) {
  if (!shard || n < SH_INDEX_BYTES) return 0;
  const uint8_t *index = shard + (n - SH_INDEX_BYTES);
  const uint8_t *t = index + (uint64_t)SH_N * 16u;
  const uint32_t want = (uint32_t)t[0] | (uint32_t)t[1] << 8 | (uint32_t)t[2] << 16 | (uint32_t)t[3] << 24;
  return r3do_crc32c(index, (size_t)SH_N * 16u) == want;
}

/* shardio.c:157-172: 0 present (*off, *len), 1 missing, -1 invalid. */
int r3do_shard_chunk_location(const uint8_t *shard, uint64_t n, uint32_t cz, uint32_t cy, uint32_t cx,
                              uint64_t *off, uint64_t *len) {
  if (cz >= SH_GRID || cy >= SH_GRID || cx >= SH_GRID || n < SH_INDEX_BYTES) return -1;
  const uint8_t *e = shard + (n - SH_INDEX_BYTES) + (((uint64_t)cz * SH_GRID + cy) * SH_GRID + cx) * 16u;
  const uint64_t o = le64(e), l = le64(e + 8);
  if (o == UINT64_MAX && l == UINT64_MAX) return 1;
  if (o == UINT64_MAX || l == UINT64_MAX) return -1;
  const uint64_t payload = n - SH_INDEX_BYTES;
  if (o > payload || l > payload - o) return -1;
  *off = o;
  *len = l;
  return 0;
}

/* One chunk: 1 decoded, 0 missing (zeros), -1 corrupt (zeros, as
 * r3d_shard_chunk_decode's -1, shardio.c:178-197), -2 levels beyond int16
 * (zeros; the reference decodes these, the int16 level formats cannot). */
static int chunk_decode(const r3do_shard *sh, int valid, uint32_t cz, uint32_t cy, uint32_t cx, uint8_t out[4096]) {
  memset(out, 0, 4096);
  if (!sh) return 0;        /* absent shard: zeros */
  if (!valid) return -1;    /* corrupt shard */
  uint64_t off = 0, len = 0;
  const int loc = r3do_shard_chunk_location(sh->bytes, sh->n, cz, cy, cx, &off, &len);
  if (loc > 0) return 0;
  if (loc < 0) return -1;
  r3do_blob_hdr h;
  int16_t lvl[4096];
  uint16_t cp[4096];
  float cd[4096], step[46];
  uint32_t nc = 0;
  const int st = r3do_parse_blob(sh->bytes + off, (size_t)len, 1, &h, lvl, &nc, cp, cd);
  if (st == 0) return -1;
  if (st < 0) return -2;
  r3do_step_table_qstep(h.qstep, step);
  const r3do_chunk_hdr ch = {h.vmin, h.vspan, h.dcq, 0};
  if (!r3do_decode_chunk(lvl, &ch, step, cp, cd, nc, out, NULL)) {
    memset(out, 0, 4096);
    return -1;
  }
  return 1;
}

/* r3d_shard_decode_region (shardio.c:207-306): dst (x fastest, dz x dy x dx)
 * is zero-filled, the region is clamped to the volume, and every chunk of the
 * clamped region is located in its shard and decoded; the part of it inside
 * the region is copied. shards lists the present shards (by shard
 * coordinates); an unlisted shard is absent (zeros). Returns 0, -1 when any
 * chunk or shard was corrupt, -1 without touching dst for a size that
 * overflows. counts (optional) accumulates decoded, missing, corrupt and
 * beyond-int16 chunks; chunk_status (optional) receives each chunk's status in
 * the clamped chunk box (x fastest). */
int r3do_shard_decode_region(const r3do_shard *shards, uint32_t nshards, uint64_t nz, uint64_t ny, uint64_t nx,
                             uint64_t z0, uint64_t y0, uint64_t x0, uint32_t dz, uint32_t dy, uint32_t dx,
                             uint8_t *dst, uint32_t counts[4], int32_t *chunk_status, int nthreads) {
  if (!dst || dz == 0 || dy == 0 || dx == 0 || (size_t)dy > SIZE_MAX / (size_t)dx ||
      (size_t)dz > SIZE_MAX / ((size_t)dy * (size_t)dx))
    return -1;
  memset(dst, 0, (size_t)dz * dy * dx);
  if (z0 >= nz || y0 >= ny || x0 >= nx) return 0;
  const uint64_t z1 = dz < nz - z0 ? z0 + dz : nz;
  const uint64_t y1 = dy < ny - y0 ? y0 + dy : ny;
  const uint64_t x1 = dx < nx - x0 ? x0 + dx : nx;
  const uint64_t c0z = z0 / SH_CHUNK, c0y = y0 / SH_CHUNK, c0x = x0 / SH_CHUNK;
  const uint64_t ncz = (z1 - 1) / SH_CHUNK - c0z + 1, ncy = (y1 - 1) / SH_CHUNK - c0y + 1,
                 ncx = (x1 - 1) / SH_CHUNK - c0x + 1;
  const uint64_t total = ncz * ncy * ncx;
  /* shard validity once per listed shard */
  int valid_buf[64];
  int *valid = nshards <= 64 ? valid_buf : (int *)__builtin_alloca(nshards * sizeof(int));
  for (uint32_t s = 0; s < nshards; s++) valid[s] = r3do_shard_valid(shards[s].bytes, shards[s].n);
  uint32_t cnt[4] = {0, 0, 0, 0};
  if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads) reduction(+ : cnt[:4])
  for (uint64_t i = 0; i < total; i++) {
    const uint64_t cz = c0z + i / (ncy * ncx), cy = c0y + (i / ncx) % ncy, cx = c0x + i % ncx;
    const uint64_t sz = cz / SH_GRID, sy = cy / SH_GRID, sx = cx / SH_GRID;
    const r3do_shard *sh = NULL;
    int v = 0;
    for (uint32_t s = 0; s < nshards; s++)
      if (shards[s].sz == sz && shards[s].sy == sy && shards[s].sx == sx) {
        sh = &shards[s];
        v = valid[s];
        break;
      }
    uint8_t chunk[4096];
    const int st = chunk_decode(sh, v, (uint32_t)(cz % SH_GRID), (uint32_t)(cy % SH_GRID), (uint32_t)(cx % SH_GRID),
                                chunk);
    cnt[st == 1 ? 0 : st == 0 ? 1 : st == -1 ? 2 : 3]++;
    if (chunk_status) chunk_status[i] = st;
    const uint64_t wz0 = cz * SH_CHUNK, wy0 = cy * SH_CHUNK, wx0 = cx * SH_CHUNK;
    const uint64_t iz0 = wz0 > z0 ? wz0 : z0, iy0 = wy0 > y0 ? wy0 : y0, ix0 = wx0 > x0 ? wx0 : x0;
    const uint64_t iz1 = wz0 + SH_CHUNK < z1 ? wz0 + SH_CHUNK : z1, iy1 = wy0 + SH_CHUNK < y1 ? wy0 + SH_CHUNK : y1,
                   ix1 = wx0 + SH_CHUNK < x1 ? wx0 + SH_CHUNK : x1;
    for (uint64_t z = iz0; z < iz1; z++)
      for (uint64_t y = iy0; y < iy1; y++)
        memcpy(dst + ((z - z0) * dy + (y - y0)) * dx + (ix0 - x0), chunk + ((z - wz0) * SH_CHUNK + (y - wy0)) * SH_CHUNK + (ix0 - wx0),
               ix1 - ix0);
  }
  if (counts)
    for (int k = 0; k < 4; k++) counts[k] += cnt[k];
  return cnt[2] ? -1 : 0;
}
