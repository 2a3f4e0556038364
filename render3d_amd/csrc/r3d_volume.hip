// r3d_volume.hip -- bricked mip pyramid, occupancy and the synthetic sheet
// volume.
//
// HBM layout. Each level is a grid of 16^3 bricks; a brick is stored as
// 17x17x17 bytes (x fastest): the 16^3 payload plus a one-voxel apron on the
// +x, +y and +z faces holding the neighbouring brick's first voxel (or, at the
// volume's edge and in partial bricks, the clamped edge voxel). A trilinear
// footprint starting at payload voxel (x0,y0,z0) therefore lies inside one
// brick at the fixed offsets {0, 1, 17, 18, 289, 290, 306, 307}, which is the
// clamp-to-edge sampler of src/vk/vkbackend.c:1058-1069 without per-tap
// clamping. A 16^3 brick is also exactly one 3ddct chunk, so the decoder
// writes level 0 in place (strides 17 / 289).
//
// Levels: dims(l) = max(dims(0) >> l, 1); level l+1 is the rounded 2x box of
// level l, (sum of 8 + 4) / 8 with the source index clamped to the level's
// last voxel (tools/lodpack/main.c:216-227). The reference builds its chain
// with a linear blit whose rounding is implementation-defined
// (src/vk/vkbackend.c:944-983); the exact integer box replaces it.
//
// Occupancy: max over each 8^3 block of level 0, then a 3^3 max dilation
// (src/vk/vkbackend.c:794-870; occmax.comp, occdilate.comp).
#include <new>
#include <stdlib.h>
#include <string.h>

#include "r3d_dmath.h"
#include "r3d_internal.h"

namespace r3d {

struct LevelGeom {
  uint8_t *base;  // level's first brick
  uint32_t nx, ny, nz;
  uint32_t bx, by, bz;
};

__device__ __forceinline__ size_t payload_index(const LevelGeom &L, uint32_t x, uint32_t y, uint32_t z) {
  const size_t b = ((size_t)(z >> 4) * L.by + (y >> 4)) * L.bx + (x >> 4);
  return b * R3D_BRICK_BYTES + ((z & 15u) * 17u + (y & 15u)) * 17u + (x & 15u);
}

// Level 0 from a source of voxels, one workgroup per brick: the 17^3 brick
// (payload, apron, padding = the voxel at the clamped global position) is
// assembled in LDS and written with 16-byte stores; the same pass emits the
// brick's eight raw 8^3 occupancy maxima and its octant of level 1 (the
// rounded 2x box of its payload). SRC_CHUNKS reads flat 4096-byte chunks in
// brick order (dims multiples of 16); SRC_LINEAR reads [z][y][x] voxels.
enum { SRC_CHUNKS = 0, SRC_LINEAR = 1 };

template <int SRC>
__device__ __forceinline__ uint32_t src_voxel(const uint8_t *src, const LevelGeom &L, uint32_t x, uint32_t y,
                                              uint32_t z) {
  if (SRC == SRC_CHUNKS)
    return src[(((size_t)(z >> 4) * L.by + (y >> 4)) * L.bx + (x >> 4)) * 4096u + ((z & 15u) * 16u + (y & 15u)) * 16u +
               (x & 15u)];
  return src[((size_t)z * L.ny + y) * L.nx + x];
}

template <int SRC>
__global__ __launch_bounds__(256) void k_brickify(const uint8_t *src, LevelGeom L0, LevelGeom L1, int do_l1,
                                                  uint8_t *raw, uint32_t ox, uint32_t oy, uint32_t oz) {
  __shared__ __attribute__((aligned(16))) uint8_t tile[R3D_BRICK_BYTES];
  __shared__ uint32_t m8[8];
  const uint32_t b = blockIdx.x;
  const uint32_t bxi = b % L0.bx, byi = (b / L0.bx) % L0.by, bzi = b / (L0.bx * L0.by);
  const uint32_t t = threadIdx.x;
  const uint32_t lz = t >> 4, ly = t & 15u;
  const bool full = bxi * 16u + 16u <= L0.nx && byi * 16u + 16u <= L0.ny && bzi * 16u + 16u <= L0.nz;
  const bool vec = full && (SRC == SRC_CHUNKS || (L0.nx & 15u) == 0u);
  if (t < 8u) m8[t] = 0u;
  if (t < R3D_BRICK_BYTES - R3D_BRICK_CONTENT) tile[R3D_BRICK_CONTENT + t] = 0;
  uint32_t row_lo = 0, row_hi = 0;  // maxima of the row's x < 8 and x >= 8 halves (vec rows)
  if (vec) {  // own payload row as one 16-byte load
    const uint8_t *row = SRC == SRC_CHUNKS
                             ? src + (size_t)b * 4096u + t * 16u
                             : src + ((size_t)(bzi * 16u + lz) * L0.ny + (byi * 16u + ly)) * L0.nx + bxi * 16u;
    const uint4 q = *(const uint4 *)row;
    const uint32_t w[4] = {q.x, q.y, q.z, q.w};
    uint8_t *d = tile + (lz * 17u + ly) * 17u;
#pragma unroll
    for (int i = 0; i < 16; i++) {
      const uint32_t v = (w[i >> 2] >> ((i & 3) * 8)) & 255u;
      d[i] = (uint8_t)v;
      if (i < 8) row_lo = v > row_lo ? v : row_lo;
      else row_hi = v > row_hi ? v : row_hi;
    }
  }
  if (vec) {
    // apron only: 256 z-face + 272 y-face + 289 x-face bytes, enumerated directly
    auto put = [&](uint32_t px, uint32_t py, uint32_t pz) {
      uint32_t x = bxi * 16u + px, y = byi * 16u + py, z = bzi * 16u + pz;
      x = x < L0.nx ? x : L0.nx - 1u;
      y = y < L0.ny ? y : L0.ny - 1u;
      z = z < L0.nz ? z : L0.nz - 1u;
      tile[(pz * 17u + py) * 17u + px] = (uint8_t)src_voxel<SRC>(src, L0, x, y, z);
    };
    const uint32_t a = t >> 4, c = t & 15u;
    put(c, a, 16u);   // z = 16, y < 16, x < 16
    put(c, 16u, a);   // y = 16, z < 16, x < 16
    put(16u, c, a);   // x = 16, z < 16, y < 16
    if (t < 16u) put(t, 16u, 16u);                 // y = z = 16, x < 16
    else if (t < 33u) put(16u, t - 16u, 16u);      // x = 16, z = 16, y = 0..16
    else if (t < 49u) put(16u, 16u, t - 33u);      // x = 16, y = 16, z < 16
  } else {
    for (uint32_t p = t; p < R3D_BRICK_CONTENT; p += 256u) {
      const uint32_t pz = p / 289u, py = (p / 17u) % 17u, px = p % 17u;
      uint32_t x = bxi * 16u + px, y = byi * 16u + py, z = bzi * 16u + pz;
      x = x < L0.nx ? x : L0.nx - 1u;
      y = y < L0.ny ? y : L0.ny - 1u;
      z = z < L0.nz ? z : L0.nz - 1u;
      tile[p] = (uint8_t)src_voxel<SRC>(src, L0, x, y, z);
    }
  }
  __syncthreads();
  uint4 *dst = (uint4 *)(L0.base + (size_t)b * R3D_BRICK_BYTES);
  for (uint32_t i = t; i < R3D_BRICK_BYTES / 16u; i += 256u) dst[i] = ((const uint4 *)tile)[i];
  // raw occupancy: payload voxels inside the level only (a vec row's maxima
  // were taken from its registers)
  {
    const uint32_t gy = byi * 16u + ly, gz = bzi * 16u + lz;
    if (gy < L0.ny && gz < L0.nz) {
      uint32_t lo = row_lo, hi = row_hi;
      if (!vec) {
        const uint8_t *row = tile + (lz * 17u + ly) * 17u;
        const uint32_t xn = L0.nx - bxi * 16u < 16u ? L0.nx - bxi * 16u : 16u;
        for (uint32_t x = 0; x < xn; x++) {
          const uint32_t v = row[x];
          if (x < 8u) lo = v > lo ? v : lo;
          else hi = v > hi ? v : hi;
        }
      }
      const uint32_t q = ((lz >> 3) * 2u + (ly >> 3)) * 2u;
      atomicMax(&m8[q], lo);
      atomicMax(&m8[q + 1u], hi);
    }
  }
  // level-1 octant: 8^3 voxels, two per lane (all 32 tile reads independent),
  // sources clamped through the tile
  if (do_l1 && (bxi >> 1) < L1.bx && (byi >> 1) < L1.by && (bzi >> 1) < L1.bz) {
    uint8_t *doct = L1.base + ((size_t)((bzi >> 1) * L1.by + (byi >> 1)) * L1.bx + (bxi >> 1)) * R3D_BRICK_BYTES +
                    (((bzi & 1u) * 8u) * 17u + (byi & 1u) * 8u) * 17u + (bxi & 1u) * 8u;
    uint32_t sum[2];
#pragma unroll
    for (int k = 0; k < 2; k++) {
      const uint32_t o = t + 256u * (uint32_t)k;
      const uint32_t x1 = o & 7u, y1 = (o >> 3) & 7u, z1 = o >> 6;
      const uint8_t *r00 = tile + ((2u * z1) * 17u + 2u * y1) * 17u + 2u * x1;
      uint32_t acc = 4u;
#pragma unroll
      for (uint32_t dz = 0; dz < 2; dz++)
#pragma unroll
        for (uint32_t dy = 0; dy < 2; dy++) {
          const uint8_t *r = r00 + (dz * 17u + dy) * 17u;
          acc += (uint32_t)r[0] + (uint32_t)r[1];
        }
      sum[k] = acc;
    }
#pragma unroll
    for (int k = 0; k < 2; k++) {
      const uint32_t o = t + 256u * (uint32_t)k;
      const uint32_t x1 = o & 7u, y1 = (o >> 3) & 7u, z1 = o >> 6;
      doct[(z1 * 17u + y1) * 17u + x1] = (uint8_t)(sum[k] / 8u);
    }
  }
  __syncthreads();
  if (t < 8u) {
    const uint32_t X = bxi * 2u + (t & 1u), Y = byi * 2u + ((t >> 1) & 1u), Z = bzi * 2u + (t >> 2);
    if (X < ox && Y < oy && Z < oz) raw[((size_t)Z * oy + Y) * ox + X] = (uint8_t)m8[t];
  }
}

// Level 0 from chunks, four bricks along x per workgroup (L0.bx a multiple of
// 4, dims multiples of 16, so every brick is full). Each lane loads its row
// of all four payloads (16 bytes each) up front; the z = 16 and y = 16 apron
// faces and their edge are 16-byte rows of the neighbour chunks (or of the
// brick's own last layer at the volume's end, as the clamp gives); only the
// group's last brick reads its x = 16 face from the next chunk, and bricks
// 0..2 copy theirs from the next brick's x = 0 plane in LDS. The four bricks
// are contiguous in the level, so they go out as one run of 16-byte stores.
// Same bytes, occupancy maxima and level-1 octants as k_brickify: a lane has
// four times the bytes in flight, and the +x chunk's strided column is read
// once per four bricks instead of once per brick.
constexpr uint32_t BK_GROUP = 4u;

__global__ __launch_bounds__(256) void k_brickify4(const uint8_t *src, LevelGeom L0, LevelGeom L1, int do_l1,
                                                   uint8_t *raw, uint32_t ox, uint32_t oy, uint32_t oz) {
  __shared__ __attribute__((aligned(16))) uint8_t tile[BK_GROUP * R3D_BRICK_BYTES];
  __shared__ uint32_t m8[BK_GROUP * 8u];
  const uint32_t ng = L0.bx / BK_GROUP;
  const uint32_t b = blockIdx.x;
  const uint32_t bx0 = (b % ng) * BK_GROUP, byi = (b / ng) % L0.by, bzi = b / (ng * L0.by);
  const uint32_t t = threadIdx.x, lz = t >> 4, ly = t & 15u;
  const size_t b0 = ((size_t)bzi * L0.by + byi) * L0.bx + bx0;  // the group's first brick (= chunk)
  const bool yn = byi + 1u < L0.by, zn = bzi + 1u < L0.bz;
  const size_t sy = (size_t)L0.bx * 4096u, sz = (size_t)L0.bx * L0.by * 4096u;  // chunk strides in bytes
  const uint8_t *c0 = src + b0 * 4096u;
  uint4 P[BK_GROUP];
#pragma unroll
  for (uint32_t k = 0; k < BK_GROUP; k++) P[k] = *(const uint4 *)(c0 + k * 4096u + t * 16u);
  // lanes 0..63: the z = 16 face rows (brick fk, y = fi); 64..127: the y = 16
  // face rows (brick fk, z = fi); 128..131: the y = z = 16 edge rows
  const uint32_t fk = (t >> 4) & 3u, fi = t & 15u;
  uint4 F = make_uint4(0, 0, 0, 0);
  uint32_t fo = 0;  // LDS offset of the row within the group's tiles
  if (t < 64u) {
    F = *(const uint4 *)(c0 + fk * 4096u + (zn ? sz : 15u * 256u) + fi * 16u);
    fo = fk * R3D_BRICK_BYTES + (16u * 17u + fi) * 17u;
  } else if (t < 128u) {
    F = *(const uint4 *)(c0 + fk * 4096u + (yn ? sy : 15u * 16u) + fi * 256u);
    fo = fk * R3D_BRICK_BYTES + (fi * 17u + 16u) * 17u;
  } else if (t < 128u + BK_GROUP) {
    const uint32_t k = t - 128u;
    F = *(const uint4 *)(c0 + k * 4096u + (yn ? sy : 15u * 16u) + (zn ? sz : 15u * 256u));
    fo = k * R3D_BRICK_BYTES + (16u * 17u + 16u) * 17u;
  }
  // the last brick's x = 16 face (17 x 17, edges included) from the next chunk's
  // x = 0 column, clamped to the volume
  uint32_t xv[2] = {0u, 0u};
  {
    const uint32_t X = min(16u * (bx0 + BK_GROUP), L0.nx - 1u);
#pragma unroll
    for (int h = 0; h < 2; h++) {
      const uint32_t i = t + 256u * (uint32_t)h;
      if (i < 289u) {
        const uint32_t y = min(byi * 16u + i % 17u, L0.ny - 1u), z = min(bzi * 16u + i / 17u, L0.nz - 1u);
        xv[h] = src_voxel<SRC_CHUNKS>(src, L0, X, y, z);
      }
    }
  }
  if (t < BK_GROUP * 8u) m8[t] = 0u;
  if (t < BK_GROUP * (R3D_BRICK_BYTES - R3D_BRICK_CONTENT)) {
    const uint32_t pad = R3D_BRICK_BYTES - R3D_BRICK_CONTENT;
    tile[(t / pad) * R3D_BRICK_BYTES + R3D_BRICK_CONTENT + t % pad] = 0;
  }
  uint32_t row_lo[BK_GROUP], row_hi[BK_GROUP];  // maxima of the row's x < 8 and x >= 8 halves
#pragma unroll
  for (uint32_t k = 0; k < BK_GROUP; k++) {
    const uint32_t w[4] = {P[k].x, P[k].y, P[k].z, P[k].w};
    uint8_t *d = tile + k * R3D_BRICK_BYTES + (lz * 17u + ly) * 17u;
    row_lo[k] = 0u;
    row_hi[k] = 0u;
#pragma unroll
    for (int i = 0; i < 16; i++) {
      const uint32_t v = (w[i >> 2] >> ((i & 3) * 8)) & 255u;
      d[i] = (uint8_t)v;
      if (i < 8) row_lo[k] = v > row_lo[k] ? v : row_lo[k];
      else row_hi[k] = v > row_hi[k] ? v : row_hi[k];
    }
  }
  if (t < 128u + BK_GROUP) {
    const uint32_t w[4] = {F.x, F.y, F.z, F.w};
#pragma unroll
    for (int i = 0; i < 16; i++) tile[fo + i] = (uint8_t)((w[i >> 2] >> ((i & 3) * 8)) & 255u);
  }
#pragma unroll
  for (int h = 0; h < 2; h++) {
    const uint32_t i = t + 256u * (uint32_t)h;
    if (i < 289u) tile[(BK_GROUP - 1u) * R3D_BRICK_BYTES + i * 17u + 16u] = (uint8_t)xv[h];
  }
  __syncthreads();
  // bricks 0..2: x = 16 face = the next brick's x = 0 plane (its y and z
  // aprons included, clamped the same way)
  for (uint32_t i = t; i < (BK_GROUP - 1u) * 289u; i += 256u) {
    const uint32_t k = i / 289u, j = i % 289u;
    tile[k * R3D_BRICK_BYTES + j * 17u + 16u] = tile[(k + 1u) * R3D_BRICK_BYTES + j * 17u];
  }
  {
    const uint32_t q = ((lz >> 3) * 2u + (ly >> 3)) * 2u;
#pragma unroll
    for (uint32_t k = 0; k < BK_GROUP; k++) {
      atomicMax(&m8[k * 8u + q], row_lo[k]);
      atomicMax(&m8[k * 8u + q + 1u], row_hi[k]);
    }
  }
  __syncthreads();
  uint4 *dst = (uint4 *)(L0.base + b0 * R3D_BRICK_BYTES);
  for (uint32_t i = t; i < BK_GROUP * R3D_BRICK_BYTES / 16u; i += 256u) dst[i] = ((const uint4 *)tile)[i];
  // level-1 octants: 8^3 voxels per brick, two per lane and brick
  if (do_l1 && (byi >> 1) < L1.by && (bzi >> 1) < L1.bz) {
#pragma unroll
    for (uint32_t k = 0; k < BK_GROUP; k++) {
      const uint32_t bxi = bx0 + k;
      if ((bxi >> 1) >= L1.bx) continue;
      uint8_t *doct = L1.base +
                      ((size_t)((bzi >> 1) * L1.by + (byi >> 1)) * L1.bx + (bxi >> 1)) * R3D_BRICK_BYTES +
                      (((bzi & 1u) * 8u) * 17u + (byi & 1u) * 8u) * 17u + (bxi & 1u) * 8u;
      const uint8_t *tk = tile + k * R3D_BRICK_BYTES;
      uint32_t sum[2];
#pragma unroll
      for (int h = 0; h < 2; h++) {
        const uint32_t o = t + 256u * (uint32_t)h;
        const uint32_t x1 = o & 7u, y1 = (o >> 3) & 7u, z1 = o >> 6;
        const uint8_t *r00 = tk + ((2u * z1) * 17u + 2u * y1) * 17u + 2u * x1;
        uint32_t acc = 4u;
#pragma unroll
        for (uint32_t dz = 0; dz < 2; dz++)
#pragma unroll
          for (uint32_t dy = 0; dy < 2; dy++) {
            const uint8_t *r = r00 + (dz * 17u + dy) * 17u;
            acc += (uint32_t)r[0] + (uint32_t)r[1];
          }
        sum[h] = acc;
      }
#pragma unroll
      for (int h = 0; h < 2; h++) {
        const uint32_t o = t + 256u * (uint32_t)h;
        const uint32_t x1 = o & 7u, y1 = (o >> 3) & 7u, z1 = o >> 6;
        doct[(z1 * 17u + y1) * 17u + x1] = (uint8_t)(sum[h] / 8u);
      }
    }
  }
  if (t < BK_GROUP * 8u) {
    const uint32_t k = t >> 3, j = t & 7u;
    const uint32_t X = (bx0 + k) * 2u + (j & 1u), Y = byi * 2u + ((j >> 1) & 1u), Z = bzi * 2u + (j >> 2);
    if (X < ox && Y < oy && Z < oz) raw[((size_t)Z * oy + Y) * ox + X] = (uint8_t)m8[t];
  }
}

// One workgroup per brick. Apron and out-of-level bytes copy the payload
// voxel at the clamped global position; payload bytes inside the level are
// only read, so concurrent bricks never race. A full brick's 817 apron bytes
// (x = 16 face 256, y = 16 face 272, z = 16 face 289) are enumerated
// directly, up to four independent items per lane with every load issued
// before any store (a lane walking a 17-byte apron row made a serial chain of
// 17 dependent memory round trips); partial bricks walk their 4913 bytes.
__global__ __launch_bounds__(256) void k_apron(LevelGeom L) {
  const uint32_t b = xcd_band(blockIdx.x, gridDim.x);  // the +x neighbour brick on the same XCD
  const uint32_t bxi = b % L.bx, byi = (b / L.bx) % L.by, bzi = b / (L.bx * L.by);
  const uint32_t t = threadIdx.x;
  const uint32_t xf = min(16u, L.nx - bxi * 16u), yf = min(16u, L.ny - byi * 16u), zf = min(16u, L.nz - bzi * 16u);
  uint8_t *dst = L.base + (size_t)b * R3D_BRICK_BYTES;
  const auto src = [&](uint32_t lx, uint32_t ly, uint32_t lz) {
    const uint32_t gx = min(bxi * 16u + lx, L.nx - 1u), gy = min(byi * 16u + ly, L.ny - 1u),
                   gz = min(bzi * 16u + lz, L.nz - 1u);
    return L.base[payload_index(L, gx, gy, gz)];
  };
  if (xf == 16u && yf == 16u && zf == 16u) {
    uint32_t pos[4];
    uint8_t val[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const uint32_t i = t + 256u * (uint32_t)k;
      uint32_t lx = 16u, ly = i & 15u, lz = i >> 4;  // x face
      if (i >= 256u && i < 528u) {                  // y face, lz < 16
        const uint32_t j = i - 256u;
        lx = j % 17u;
        ly = 16u;
        lz = j / 17u;
      } else if (i >= 528u) {  // z face
        const uint32_t j = min(i - 528u, 288u);
        lx = j % 17u;
        ly = j / 17u;
        lz = 16u;
      }
      pos[k] = (lz * 17u + ly) * 17u + lx;
      val[k] = i < 817u ? src(lx, ly, lz) : 0u;
    }
#pragma unroll
    for (int k = 0; k < 4; k++)
      if (t + 256u * (uint32_t)k < 817u) dst[pos[k]] = val[k];
    return;
  }
  for (uint32_t p = t; p < R3D_BRICK_CONTENT; p += 256u) {
    const uint32_t lz = p / 289u, ly = (p / 17u) % 17u, lx = p % 17u;
    if (lx < xf && ly < yf && lz < zf) continue;
    dst[p] = src(lx, ly, lz);
  }
}

// Level l+1 from level l, one workgroup per destination brick; every byte
// (payload, apron) is the rounded 2x box at its clamped global position, the
// eight sources clamped to the source level. Four destination bytes per
// round with all 32 loads in flight.
__global__ __launch_bounds__(256) void k_downsample(LevelGeom S, LevelGeom D) {
  const uint32_t b = blockIdx.x;
  const uint32_t bxi = b % D.bx, byi = (b / D.bx) % D.by, bzi = b / (D.bx * D.by);
  const uint32_t t = threadIdx.x;
  uint8_t *dst = D.base + (size_t)b * R3D_BRICK_BYTES;
  for (uint32_t p0 = t; p0 < R3D_BRICK_CONTENT; p0 += 1024u) {
    uint32_t sum[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const uint32_t p = min(p0 + 256u * (uint32_t)k, R3D_BRICK_CONTENT - 1u);
      const uint32_t lz = p / 289u, ly = (p / 17u) % 17u, lx = p % 17u;
      const uint32_t x = min(bxi * 16u + lx, D.nx - 1u), y = min(byi * 16u + ly, D.ny - 1u),
                     z = min(bzi * 16u + lz, D.nz - 1u);
      // The box's sources 2x, 2x + 1 (per axis) lie in one source brick (2x is
      // even) and 2x <= S.n - 1 (level dims are floor halvings); 2x + 1 is
      // clamped to 2x only on a one-voxel source axis. So one index and the
      // payload strides (1, 17, 289), zero where clamped, give all eight.
      const uint8_t *s0 = S.base + payload_index(S, 2u * x, 2u * y, 2u * z);
      const uint32_t oy = 2u * y + 1u < S.ny ? 17u : 0u, oz = 2u * z + 1u < S.nz ? 289u : 0u;
      if (S.nx >= 2u) {  // wave-uniform: 2x + 1 < S.nx, so each x pair is one 2-byte load
        uint16_t q[4];
        __builtin_memcpy(&q[0], s0, 2);
        __builtin_memcpy(&q[1], s0 + oy, 2);
        __builtin_memcpy(&q[2], s0 + oz, 2);
        __builtin_memcpy(&q[3], s0 + oy + oz, 2);
        uint32_t acc = 4u;
#pragma unroll
        for (int i = 0; i < 4; i++) acc += (uint32_t)(q[i] & 255u) + (uint32_t)(q[i] >> 8);
        sum[k] = acc;
      } else {
        sum[k] = 4u + 2u * ((uint32_t)s0[0] + s0[oy] + s0[oz] + s0[oy + oz]);
      }
    }
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const uint32_t p = p0 + 256u * (uint32_t)k;
      if (p < R3D_BRICK_CONTENT) dst[p] = (uint8_t)(sum[k] / 8u);
    }
  }
}

// Raw occupancy of level 0, one workgroup per brick (its 2x2x2 blocks of 8^3),
// one lane per payload row.
__global__ __launch_bounds__(256) void k_occ_raw(LevelGeom L, uint8_t *raw, uint32_t ox, uint32_t oy, uint32_t oz) {
  __shared__ uint32_t m8[8];
  const uint32_t b = blockIdx.x;
  const uint32_t bxi = b % L.bx, byi = (b / L.bx) % L.by, bzi = b / (L.bx * L.by);
  const uint32_t t = threadIdx.x;
  if (t < 8u) m8[t] = 0u;
  __syncthreads();
  const uint32_t lz = t >> 4, ly = t & 15u;
  const uint32_t gy = byi * 16u + ly, gz = bzi * 16u + lz;
  if (gy < L.ny && gz < L.nz) {
    const uint8_t *row = L.base + (size_t)b * R3D_BRICK_BYTES + (lz * 17u + ly) * 17u;
    const uint32_t xn = L.nx - bxi * 16u < 16u ? L.nx - bxi * 16u : 16u;
    uint32_t lo = 0, hi = 0;
    for (uint32_t x = 0; x < xn; x++) {
      const uint32_t v = row[x];
      if (x < 8u) lo = v > lo ? v : lo;
      else hi = v > hi ? v : hi;
    }
    const uint32_t q = ((lz >> 3) * 2u + (ly >> 3)) * 2u;
    atomicMax(&m8[q], lo);
    atomicMax(&m8[q + 1u], hi);
  }
  __syncthreads();
  if (t < 8u) {
    const uint32_t X = bxi * 2u + (t & 1u), Y = byi * 2u + ((t >> 1) & 1u), Z = bzi * 2u + (t >> 2);
    if (X < ox && Y < oy && Z < oz) raw[((size_t)Z * oy + Y) * ox + X] = (uint8_t)m8[t];
  }
}

__global__ void k_occ_dilate(const uint8_t *raw, uint8_t *occ, uint32_t ox, uint32_t oy, uint32_t oz) {
  const uint64_t n = (uint64_t)ox * oy * oz;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const int x = (int)(i % ox), y = (int)((i / ox) % oy), z = (int)(i / ((uint64_t)ox * oy));
    uint32_t m = 0;
    for (int dz = -1; dz <= 1; dz++)
      for (int dy = -1; dy <= 1; dy++)
        for (int dx = -1; dx <= 1; dx++) {
          const int zz = z + dz, yy = y + dy, xx = x + dx;
          if (zz < 0 || yy < 0 || xx < 0 || zz >= (int)oz || yy >= (int)oy || xx >= (int)ox) continue;
          const uint32_t v = raw[((size_t)zz * oy + yy) * ox + xx];
          m = v > m ? v : m;
        }
    occ[i] = (uint8_t)m;
  }
}

__global__ void k_level_to_linear(LevelGeom L, uint8_t *dst) {
  const uint64_t n = (uint64_t)L.nx * L.ny * L.nz;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t x = (uint32_t)(i % L.nx), y = (uint32_t)((i / L.nx) % L.ny),
                   z = (uint32_t)(i / ((uint64_t)L.nx * L.ny));
    dst[i] = L.base[payload_index(L, x, y, z)];
  }
}

__global__ void k_chunk_offsets(uint64_t *off, uint64_t level_off, uint32_t bx, uint32_t by, uint32_t bz) {
  const uint64_t n = (uint64_t)bx * by * bz;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    off[i] = level_off + i * R3D_BRICK_BYTES;
}

// The sheet volume; same pinned sequence as the host generator.
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__global__ void k_synth_sheets(uint8_t *vol, uint32_t nx, uint32_t ny, uint32_t nz, uint32_t nsheets,
                               uint64_t seed) {
  __shared__ int32_t off[256];
  for (uint32_t k = threadIdx.x; k < nsheets; k += blockDim.x)
    off[k] = (int32_t)(mix64(seed ^ (0x5EED0000ull + k)) % 61u) - 30;
  __syncthreads();
  const float cx = (float)nx * 0.5f, cy = (float)ny * 0.5f;
  const float LN2 = 0x1.62e430p-1f;
  const uint64_t n = (uint64_t)nx * ny * nz;
  for (uint64_t idx = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; idx < n;
       idx += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t x = (uint32_t)(idx % nx), y = (uint32_t)((idx / nx) % ny),
                   z = (uint32_t)(idx / ((uint64_t)nx * ny));
    float ws, wc;
    sincos_turns((float)z / 40.0f, &ws, &wc);
    const float wave = 3.0f * ws;
    const uint64_t h = mix64(seed ^ (idx * 0xD1B54A32D192ED03ull));
    const uint32_t a = (uint32_t)h, b = (uint32_t)(h >> 32), c = (uint32_t)mix64(h);
    int32_t base = 20 + (int32_t)(c % 11u) - 5;
    const float dx = ((float)x + 0.5f) - cx, dy = ((float)y + 0.5f) - cy;
    const float r = __builtin_sqrtf(dx * dx + dy * dy);
    const float rr = r - wave;
    const float kf = __builtin_floorf((rr - 20.0f) / 9.0f + 0.5f);
    if (kf >= 0.0f && kf < (float)nsheets) {
      const int32_t k = (int32_t)kf;
      if (__builtin_fabsf(rr - (20.0f + 9.0f * (float)k)) < 1.5f) base = 160 + off[k];
    }
    const float u1 = ((float)(a >> 8) + 0.5f) * 0x1p-24f;
    const float u2 = ((float)(b >> 8) + 0.5f) * 0x1p-24f;
    float s2, c2;
    sincos_turns(u2, &s2, &c2);
    const float g = __builtin_sqrtf(-2.0f * (log2_pinned(u1) * LN2)) * c2;
    float v = __builtin_roundf((float)base + 8.0f * g);
    v = fclamp_sel(v, 0.0f, 255.0f);
    vol[idx] = (uint8_t)v;
  }
}

static LevelGeom geom(const r3d_hip_volume *v, uint32_t l) {
  LevelGeom g;
  g.base = v->d_bricks + v->level_off[l];
  g.nx = v->dims[l][0];
  g.ny = v->dims[l][1];
  g.nz = v->dims[l][2];
  g.bx = v->bricks[l][0];
  g.by = v->bricks[l][1];
  g.bz = v->bricks[l][2];
  return g;
}

static uint64_t level_bytes(const r3d_hip_volume *v, uint32_t l) {
  return (uint64_t)v->bricks[l][0] * v->bricks[l][1] * v->bricks[l][2] * R3D_BRICK_BYTES;
}

static uint32_t nbricks(const r3d_hip_volume *v, uint32_t l) {
  return v->bricks[l][0] * v->bricks[l][1] * v->bricks[l][2];
}

static dim3 grid1d(const r3d_hip_ctx *ctx, uint64_t n, uint32_t block) {
  uint64_t g = (n + block - 1) / block;
  const uint64_t cap = (uint64_t)(ctx->num_cus > 0 ? ctx->num_cus : 256) * 64u;
  return dim3((uint32_t)(g < cap ? (g ? g : 1) : cap));
}

VolDev vol_dev(const r3d_hip_volume *v) {
  VolDev d;
  d.base = v->d_bricks;
  d.levels = v->d_levels;
  d.occ = v->d_occ;
  d.nlev = v->nlev;
  uint32_t md = v->dims[0][0];
  if (v->dims[0][1] > md) md = v->dims[0][1];
  if (v->dims[0][2] > md) md = v->dims[0][2];
  d.maxdim = md;
  d.ox = v->odim[0];
  d.oy = v->odim[1];
  d.oz = v->odim[2];
  d.narrow = !v->wide && v->bytes <= ((uint64_t)1 << 32) ? 1u : 0u;
  return d;
}

}  // namespace r3d

using namespace r3d;

extern "C" int r3d_hip_volume_create(r3d_hip_ctx *ctx, const r3d_volume_desc *desc, uint32_t nlevels,
                                     r3d_hip_volume **out) {
  if (!ctx) return R3D_ERR_INVALID;
  if (!desc || !out) return r3d_fail(ctx, R3D_ERR_INVALID, "volume_create: NULL argument");
  *out = nullptr;
  if (desc->nx == 0 || desc->ny == 0 || desc->nz == 0)
    return r3d_fail(ctx, R3D_ERR_INVALID, "volume_create: empty volume");
  if (desc->nx > 65536 || desc->ny > 65536 || desc->nz > 65536)
    return r3d_fail(ctx, R3D_ERR_RANGE, "volume_create: dims above 65536");
  uint32_t maxdim = desc->nx;
  if (desc->ny > maxdim) maxdim = desc->ny;
  if (desc->nz > maxdim) maxdim = desc->nz;
  uint32_t full = 1;  // src/vk/vkbackend.c:1517-1518
  while ((maxdim >> full) >= 1 && full < R3D_MAX_LEVELS) full++;
  if (nlevels == 0) nlevels = full;
  if (nlevels > full) return r3d_fail(ctx, R3D_ERR_RANGE, "volume_create: %u levels > %u", nlevels, full);
  r3d_hip_volume *v = new (std::nothrow) r3d_hip_volume();
  if (!v) return R3D_ERR_NOMEM;
  v->ctx = ctx;
  v->desc = *desc;
  v->nlev = nlevels;
  // Kernels address the pyramid with 32-bit offsets when it fits 4 GiB;
  // R3D_WIDE_OFFSETS=1 makes this volume take the 64-bit instances a larger
  // pyramid takes (the same pixels), so they can be checked on small volumes.
  const char *wide = getenv("R3D_WIDE_OFFSETS");
  v->wide = wide && wide[0] == '1';
  uint64_t off = 0;
  LevelDev tab[R3D_MAX_LEVELS];
  memset(tab, 0, sizeof tab);
  for (uint32_t l = 0; l < nlevels; l++) {
    const uint32_t d[3] = {desc->nx, desc->ny, desc->nz};
    for (int a = 0; a < 3; a++) {
      v->dims[l][a] = (d[a] >> l) ? (d[a] >> l) : 1u;
      v->bricks[l][a] = (v->dims[l][a] + 15u) / 16u;
    }
    v->level_off[l] = off;
    off += (level_bytes(v, l) + 255u) & ~(uint64_t)255u;
    tab[l].off_lo = (uint32_t)v->level_off[l];
    tab[l].off_hi = (uint32_t)(v->level_off[l] >> 32);
    tab[l].bx = v->bricks[l][0];
    tab[l].by = v->bricks[l][1];
    tab[l].bz = v->bricks[l][2];
    tab[l].fx = (float)v->dims[l][0];
    tab[l].fy = (float)v->dims[l][1];
    tab[l].fz = (float)v->dims[l][2];
    tab[l].hx = tab[l].fx - 1.0f;
    tab[l].hy = tab[l].fy - 1.0f;
    tab[l].hz = tab[l].fz - 1.0f;
  }
  v->bytes = off;
  for (int a = 0; a < 3; a++) v->odim[a] = ((&desc->nx)[a] + 7u) / 8u;
  const uint64_t on = (uint64_t)v->odim[0] * v->odim[1] * v->odim[2];
  int rc = R3D_OK;
  if (hipSetDevice(ctx->device) != hipSuccess) rc = R3D_ERR_DEVICE;
  if (rc == R3D_OK && (hipMalloc(&v->d_bricks, v->bytes) != hipSuccess ||
                       hipMalloc(&v->d_levels, sizeof tab) != hipSuccess ||
                       hipMalloc(&v->d_occ, on) != hipSuccess || hipMalloc(&v->d_occ_raw, on) != hipSuccess))
    rc = r3d_fail(ctx, R3D_ERR_NOMEM, "volume_create: cannot allocate %llu bytes",
                  (unsigned long long)(v->bytes + 2 * on));
  if (rc == R3D_OK && hipMemcpy(v->d_levels, tab, sizeof tab, hipMemcpyHostToDevice) != hipSuccess)
    rc = R3D_ERR_DEVICE;
  if (rc != R3D_OK) {
    r3d_hip_volume_destroy(v);
    return rc;
  }
  *out = v;
  return R3D_OK;
}

extern "C" void r3d_hip_volume_destroy(r3d_hip_volume *v) {
  if (!v) return;
  if (v->ctx) hipSetDevice(v->ctx->device);
  if (v->d_bricks) hipFree(v->d_bricks);
  if (v->d_levels) hipFree(v->d_levels);
  if (v->d_occ) hipFree(v->d_occ);
  if (v->d_occ_raw) hipFree(v->d_occ_raw);
  if (v->d_flat) hipFree(v->d_flat);
  delete v;
}

extern "C" int r3d_hip_volume_info(const r3d_hip_volume *v, uint32_t *nlevels, uint32_t dims[48],
                                   uint32_t odim[3]) {
  if (!v) return R3D_ERR_INVALID;
  if (nlevels) *nlevels = v->nlev;
  if (dims)
    for (uint32_t l = 0; l < R3D_MAX_LEVELS; l++)
      for (int a = 0; a < 3; a++) dims[l * 3 + a] = l < v->nlev ? v->dims[l][a] : 0u;
  if (odim)
    for (int a = 0; a < 3; a++) odim[a] = v->odim[a];
  return R3D_OK;
}

static dim3 brick_grid(const r3d_hip_volume *v, uint32_t l) { return dim3(nbricks(v, l)); }

// Level 1 (from level 0, or only its apron when the level-0 pass already wrote
// its payload), levels 2.., and the occupancy dilation.
static int build_tail(r3d_hip_ctx *ctx, r3d_hip_volume *v, hipStream_t s, bool l1_payload_done) {
  r3d_scene_changed(ctx);  // every build ends here
  if (v->nlev > 1) {
    if (l1_payload_done)
      hipLaunchKernelGGL(k_apron, brick_grid(v, 1), dim3(256), 0, s, geom(v, 1));
    else
      hipLaunchKernelGGL(k_downsample, brick_grid(v, 1), dim3(256), 0, s, geom(v, 0), geom(v, 1));
    for (uint32_t l = 1; l + 1 < v->nlev; l++)
      hipLaunchKernelGGL(k_downsample, brick_grid(v, l + 1), dim3(256), 0, s, geom(v, l), geom(v, l + 1));
  }
  hipLaunchKernelGGL(k_occ_dilate, grid1d(ctx, (uint64_t)v->odim[0] * v->odim[1] * v->odim[2], 256), dim3(256), 0, s,
                     (const uint8_t *)v->d_occ_raw, v->d_occ, v->odim[0], v->odim[1], v->odim[2]);
  R3D_HIP_CHECK(ctx, hipGetLastError());
  return R3D_OK;
}

template <int SRC>
static int brickify_and_build(r3d_hip_ctx *ctx, r3d_hip_volume *v, const uint8_t *src, hipStream_t s) {
  const LevelGeom l1 = v->nlev > 1 ? geom(v, 1) : geom(v, 0);
  const LevelGeom l0 = geom(v, 0);
  if (SRC == SRC_CHUNKS && l0.bx % BK_GROUP == 0u)
    hipLaunchKernelGGL(k_brickify4, dim3(nbricks(v, 0) / BK_GROUP), dim3(256), 0, s, src, l0, l1,
                       v->nlev > 1 ? 1 : 0, v->d_occ_raw, v->odim[0], v->odim[1], v->odim[2]);
  else
    hipLaunchKernelGGL(k_brickify<SRC>, brick_grid(v, 0), dim3(256), 0, s, src, l0, l1, v->nlev > 1 ? 1 : 0,
                       v->d_occ_raw, v->odim[0], v->odim[1], v->odim[2]);
  R3D_HIP_CHECK(ctx, hipGetLastError());
  return build_tail(ctx, v, s, v->nlev > 1);
}

extern "C" int r3d_hip_volume_build(r3d_hip_ctx *ctx, r3d_hip_volume *v, void *stream) {
  if (!ctx) return R3D_ERR_INVALID;
  if (!v || v->ctx != ctx) return r3d_fail(ctx, R3D_ERR_INVALID, "volume_build: bad volume");
  R3D_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_apron, brick_grid(v, 0), dim3(256), 0, s, geom(v, 0));
  hipLaunchKernelGGL(k_occ_raw, brick_grid(v, 0), dim3(256), 0, s, geom(v, 0), v->d_occ_raw, v->odim[0],
                     v->odim[1], v->odim[2]);
  R3D_HIP_CHECK(ctx, hipGetLastError());
  return build_tail(ctx, v, s, false);
}

extern "C" int r3d_hip_volume_build_from_chunks(r3d_hip_ctx *ctx, r3d_hip_volume *v, const uint8_t *chunks,
                                                void *stream) {
  if (!ctx) return R3D_ERR_INVALID;
  if (!v || v->ctx != ctx || !chunks) return r3d_fail(ctx, R3D_ERR_INVALID, "build_from_chunks: bad argument");
  if (v->desc.nx % 16 || v->desc.ny % 16 || v->desc.nz % 16)
    return r3d_fail(ctx, R3D_ERR_INVALID, "build_from_chunks: level-0 dims must be multiples of 16");
  if ((uintptr_t)chunks & 15u) return r3d_fail(ctx, R3D_ERR_INVALID, "build_from_chunks: chunks not 16-byte aligned");
  R3D_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  return brickify_and_build<SRC_CHUNKS>(ctx, v, chunks, (hipStream_t)stream);
}

extern "C" int r3d_hip_volume_decode(r3d_hip_ctx *ctx, r3d_hip_volume *v, const r3d_chunk_batch *batch,
                                     const r3d_decode_opts *opts, void *stream) {
  if (!ctx) return R3D_ERR_INVALID;
  if (!v || v->ctx != ctx || !batch) return r3d_fail(ctx, R3D_ERR_INVALID, "volume_decode: bad argument");
  if (v->desc.nx % 16 || v->desc.ny % 16 || v->desc.nz % 16)
    return r3d_fail(ctx, R3D_ERR_INVALID, "volume_decode: level-0 dims must be multiples of 16");
  const uint64_t nch = (uint64_t)v->bricks[0][0] * v->bricks[0][1] * v->bricks[0][2];
  if (batch->nchunks != nch)
    return r3d_fail(ctx, R3D_ERR_INVALID, "volume_decode: %u chunks for a volume of %llu", batch->nchunks,
                    (unsigned long long)nch);
  if (opts && opts->flags) return r3d_fail(ctx, R3D_ERR_INVALID, "volume_decode: the volume takes u8 output");
  R3D_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  if (!v->d_flat && hipMalloc(&v->d_flat, nch * 4096u) != hipSuccess)
    return r3d_fail(ctx, R3D_ERR_NOMEM, "volume_decode: staging %llu bytes", (unsigned long long)(nch * 4096u));
  r3d_chunk_batch b = *batch;
  b.out = v->d_flat;
  b.out_offsets = nullptr;
  b.out_stride_y = 0;
  b.out_stride_z = 0;
  int rc = r3d_hip_decode_chunks(ctx, &b, opts, stream);
  if (rc != R3D_OK) return rc;
  return brickify_and_build<SRC_CHUNKS>(ctx, v, v->d_flat, (hipStream_t)stream);
}

extern "C" int r3d_hip_volume_upload(r3d_hip_ctx *ctx, r3d_hip_volume *v, const uint8_t *voxels,
                                     uint32_t src_kind, void *stream) {
  if (!ctx) return R3D_ERR_INVALID;
  if (!v || v->ctx != ctx || !voxels) return r3d_fail(ctx, R3D_ERR_INVALID, "volume_upload: bad argument");
  if (src_kind != R3D_SRC_HOST && src_kind != R3D_SRC_DEVICE)
    return r3d_fail(ctx, R3D_ERR_INVALID, "volume_upload: bad src_kind %u", src_kind);
  R3D_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  hipStream_t s = (hipStream_t)stream;
  const uint64_t nvox = (uint64_t)v->desc.nx * v->desc.ny * v->desc.nz;
  const uint8_t *src = voxels;
  uint8_t *tmp = nullptr;
  if (src_kind == R3D_SRC_HOST || ((uintptr_t)voxels & 15u)) {
    if (hipMalloc(&tmp, nvox) != hipSuccess)
      return r3d_fail(ctx, R3D_ERR_NOMEM, "volume_upload: staging %llu bytes", (unsigned long long)nvox);
    const hipError_t e = hipMemcpyAsync(tmp, voxels, nvox,
                                        src_kind == R3D_SRC_HOST ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice, s);
    if (e != hipSuccess) {
      hipFree(tmp);
      return r3d_fail(ctx, R3D_ERR_DEVICE, "volume_upload: copy failed");
    }
    src = tmp;
  }
  int rc = brickify_and_build<SRC_LINEAR>(ctx, v, src, s);
  if (tmp) {
    hipStreamSynchronize(s);
    hipFree(tmp);
  }
  return rc;
}

// r3d_shard_decode_region (shardio.c:269-306) of the volume's box of a shard
// store, then the volume built from it: the region decode writes the volume's
// staging buffer row by row and the brickify pass builds the bricks, aprons,
// pyramid and occupancy from there, all on the GPU.
extern "C" int r3d_hip_volume_decode_shards(r3d_hip_ctx *ctx, r3d_hip_volume *v, const r3d_shard_src *shards,
                                            uint32_t nshards, const r3d_shard_region *region, int32_t *chunk_status,
                                            uint32_t *counts, void *stream) {
  if (!ctx) return R3D_ERR_INVALID;
  if (!v || v->ctx != ctx || !region) return r3d_fail(ctx, R3D_ERR_INVALID, "volume_decode_shards: bad argument");
  if (region->dz != v->desc.nz || region->dy != v->desc.ny || region->dx != v->desc.nx)
    return r3d_fail(ctx, R3D_ERR_INVALID, "volume_decode_shards: region %ux%ux%u for a volume of %ux%ux%u",
                    region->dx, region->dy, region->dz, v->desc.nx, v->desc.ny, v->desc.nz);
  R3D_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  const uint64_t nch = (uint64_t)v->bricks[0][0] * v->bricks[0][1] * v->bricks[0][2];  // >= the voxels / 4096
  if (!v->d_flat && hipMalloc(&v->d_flat, nch * 4096u) != hipSuccess) {
    v->d_flat = nullptr;
    return r3d_fail(ctx, R3D_ERR_NOMEM, "volume_decode_shards: staging %llu bytes",
                    (unsigned long long)(nch * 4096u));
  }
  const int rc = r3d_hip_shard_decode_region(ctx, shards, nshards, region, v->d_flat, chunk_status, counts, stream);
  if (rc != R3D_OK) return rc;
  return brickify_and_build<SRC_LINEAR>(ctx, v, v->d_flat, (hipStream_t)stream);
}

extern "C" int r3d_hip_volume_chunk_targets(r3d_hip_ctx *ctx, r3d_hip_volume *v, uint64_t *out_offsets,
                                            uint8_t **base, uint32_t *stride_y, uint32_t *stride_z,
                                            void *stream) {
  if (!ctx) return R3D_ERR_INVALID;
  if (!v || v->ctx != ctx) return r3d_fail(ctx, R3D_ERR_INVALID, "chunk_targets: bad volume");
  if (v->desc.nx % 16 || v->desc.ny % 16 || v->desc.nz % 16)
    return r3d_fail(ctx, R3D_ERR_INVALID, "chunk_targets: level-0 dims must be multiples of 16");
  R3D_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  if (out_offsets) {
    const uint64_t n = (uint64_t)v->bricks[0][0] * v->bricks[0][1] * v->bricks[0][2];
    hipLaunchKernelGGL(k_chunk_offsets, grid1d(ctx, n, 256), dim3(256), 0, (hipStream_t)stream, out_offsets,
                       v->level_off[0], v->bricks[0][0], v->bricks[0][1], v->bricks[0][2]);
    R3D_HIP_CHECK(ctx, hipGetLastError());
  }
  if (base) *base = v->d_bricks;
  if (stride_y) *stride_y = R3D_BRICK_STRIDE;
  if (stride_z) *stride_z = R3D_BRICK_STRIDE * R3D_BRICK_STRIDE;
  return R3D_OK;
}

static int copy_out(r3d_hip_ctx *ctx, const uint8_t *dsrc, uint8_t *dst, uint64_t n, uint32_t kind,
                    hipStream_t s) {
  if (kind == R3D_SRC_DEVICE) {
    R3D_HIP_CHECK(ctx, hipMemcpyAsync(dst, dsrc, n, hipMemcpyDeviceToDevice, s));
  } else {
    R3D_HIP_CHECK(ctx, hipMemcpyAsync(dst, dsrc, n, hipMemcpyDeviceToHost, s));
    R3D_HIP_CHECK(ctx, hipStreamSynchronize(s));
  }
  return R3D_OK;
}

extern "C" int r3d_hip_volume_read_level(r3d_hip_ctx *ctx, const r3d_hip_volume *v, uint32_t level,
                                         uint8_t *dst, uint32_t dst_kind, void *stream) {
  if (!ctx) return R3D_ERR_INVALID;
  if (!v || v->ctx != ctx || !dst || level >= v->nlev)
    return r3d_fail(ctx, R3D_ERR_INVALID, "read_level: bad argument");
  R3D_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  hipStream_t s = (hipStream_t)stream;
  const LevelGeom g = geom(v, level);
  const uint64_t n = (uint64_t)g.nx * g.ny * g.nz;
  uint8_t *lin = nullptr;
  if (hipMalloc(&lin, n) != hipSuccess) return r3d_fail(ctx, R3D_ERR_NOMEM, "read_level: staging");
  hipLaunchKernelGGL(k_level_to_linear, grid1d(ctx, n, 256), dim3(256), 0, s, g, lin);
  int rc = copy_out(ctx, lin, dst, n, dst_kind, s);
  hipStreamSynchronize(s);
  hipFree(lin);
  return rc;
}

extern "C" int r3d_hip_volume_read_occupancy(r3d_hip_ctx *ctx, const r3d_hip_volume *v, uint8_t *dst,
                                             uint32_t dst_kind, void *stream) {
  if (!ctx) return R3D_ERR_INVALID;
  if (!v || v->ctx != ctx || !dst) return r3d_fail(ctx, R3D_ERR_INVALID, "read_occupancy: bad argument");
  R3D_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  const uint64_t n = (uint64_t)v->odim[0] * v->odim[1] * v->odim[2];
  return copy_out(ctx, v->d_occ, dst, n, dst_kind, (hipStream_t)stream);
}

extern "C" int r3d_hip_synth_sheets(r3d_hip_ctx *ctx, uint8_t *dst, uint32_t nx, uint32_t ny, uint32_t nz,
                                    uint32_t nsheets, uint64_t seed, void *stream) {
  if (!ctx) return R3D_ERR_INVALID;
  if (!dst || nx == 0 || ny == 0 || nz == 0) return r3d_fail(ctx, R3D_ERR_INVALID, "synth: bad argument");
  if (nsheets > 256) return r3d_fail(ctx, R3D_ERR_RANGE, "synth: at most 256 sheets");
  R3D_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  const uint64_t n = (uint64_t)nx * ny * nz;
  hipLaunchKernelGGL(k_synth_sheets, grid1d(ctx, n, 256), dim3(256), 0, (hipStream_t)stream, dst, nx, ny, nz,
                     nsheets, seed);
  R3D_HIP_CHECK(ctx, hipGetLastError());
  return R3D_OK;
}

extern "C" uint64_t r3d_hip_volume_brick_count(const r3d_hip_volume *v) {
  if (!v) return 0;
  uint64_t n = 0;
  for (uint32_t l = 0; l < v->nlev; l++) n += (uint64_t)v->bricks[l][0] * v->bricks[l][1] * v->bricks[l][2];
  return n;
}
