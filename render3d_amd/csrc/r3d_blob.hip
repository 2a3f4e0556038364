// r3d_blob.hip -- the 3ddct entropy stage on gfx950: blobs -> sparse levels,
// and the matching encoder for workloads.
//
// One lane per blob. The stage is the CABAC-style binary range coder of
// tools/3ddct/dct3d.c:156-405 with the framing of dct3d.c:414-433, 542-587:
// a 2-byte preamble (0xD3, dtype), then one range-coded stream holding vmin,
// vspan, qstep (float fields), dcq (signed Exp-Golomb), the corrections flag,
// the levels (EOB, banded significance contexts conditioned on the density of
// the last 16 decisions, an 11-context unary magnitude ladder with an
// Exp-Golomb escape, bypass signs) and the corrections. Every step is integer
// arithmetic, so the parse is exact by construction.
//
// The parse is bound by VALU issue, not by HBM: a blob is a serial chain of
// binary decisions (about 6,000 for a 16^3 chunk at qstep 8), so the level
// loop is a per-lane state machine making one decision per iteration (the
// lanes of a wave stay converged), renormalisation is branch-free over a
// register byte window, and the window's loads are issued 8 iterations
// before their bytes are needed.
//
// The adaptive contexts (59 per lane, 16-bit) live in LDS, one row per
// context, so lanes touching the same context hit consecutive addresses; the
// ascending-L1 scan table is staged in LDS once per workgroup.
#include <hipcub/hipcub.hpp>
#include <mutex>

#include "r3d_internal.h"

namespace r3d {

constexpr uint32_t BLOB_NT = 256;  // lanes (blobs) per workgroup
constexpr uint32_t BLK_BYTES = 32;  // the blob read block (two uint4 members)
constexpr uint32_t RC_TOP = 1u << 24;
enum : uint32_t { CTX_SIG = 0, CTX_MAG = 32, CTX_EOB = 44, CTX_FLAG = 58, NCTX = 59 };

// Ascending L1 frequency, ties by index (dct3d.c:336-351); each entry packs
// the coefficient index (bits 0-11) and its significance band (bits 12-14,
// dct3d.c:352-355).
__constant__ uint16_t c_scan[4096];
__constant__ float c_hfw[46];        // powf(1 + f, 0.65f), computed by the host's libm (dct3d.c:131-137)
// Trained priors P(bit == 0) in 1/4096 (dct3d.c:290-299), in context order.
__constant__ uint16_t c_prior[NCTX] = {
    1980, 575,  846,  743,  3916, 3707, 3515, 1463, 3969, 3753, 3527, 2313, 4002, 3826, 3635, 2819,
    4039, 3851, 3613, 2974, 3985, 3840, 3703, 3475, 4021, 3991, 3951, 3969, 4065, 4024, 2048, 2048,
    1893, 1217, 908,  728,  610,  527,  465,  416,  377,  346,  320,  297,  74,   1,    1,    1,
    1,    5,    109,  504,  772,  1063, 1458, 2326, 4095, 2048, 2048};


// ---- decode side -----------------------------------------------------------
// The stream is read through a per-lane window of up to 8 bytes held in a
// register pair (next byte in the top bits), refilled one aligned dword at a
// time. refill() issues the load for the next dword and a later merge()
// appends it, so a caller that calls them a few iterations apart hides the
// load latency; a lane that runs short in between fills synchronously.
// Aligned dword loads never cross a page the blob does not touch, and bytes
// past the end of the blob read as zero (dct3d.c:218-223).
struct RDec {
  const uint8_t *p;  // stream start
  uint32_t n;        // stream length
  uint32_t qn;       // the stream byte after the window (take() leaves it unchanged)
  uint64_t w;        // the window: bytes [qn - wn, qn), stream order from bit 63 down
  uint32_t wn;
  uint32_t stg_k, stg_skip, stg_cnt;  // a staged dword (dword stg_k of blk), pending while stg_cnt > 0
  // the aligned BLK_BYTES block the window is read from, as named members
  // (an array member would move the whole decoder state to scratch)
  uint4 b0, b1;
  uintptr_t blk_a;                    // its address (0: none yet)
  uint32_t code, range;
  uint16_t *ctx;  // this lane's contexts: ctx[i * BLOB_NT]

  // The aligned dword holding stream byte q (q < n): skip leading bytes, cnt
  // bytes from q on that lie inside the stream.
  __device__ __forceinline__ uint32_t load_at(uint32_t q, uint32_t &skip, uint32_t &cnt) const {
    const uintptr_t a = (uintptr_t)(p + q);
    skip = (uint32_t)(a & 3u);
    const uint32_t left = n - q;
    cnt = 4u - skip < left ? 4u - skip : left;
    return *(const uint32_t *)(a & ~(uintptr_t)3u);
  }
  __device__ __forceinline__ void append(uint32_t d, uint32_t skip, uint32_t cnt) {  // needs wn <= 4
    uint32_t be = __builtin_bswap32(d) << (8u * skip);
    be &= ~(0xFFFFFFFFu >> (8u * cnt - 1u) >> 1);  // keep the first cnt bytes (cnt in 1..4)
    w |= (uint64_t)be << (32u - 8u * wn);
    wn += cnt;
    qn += cnt;
  }
  __device__ __forceinline__ void merge() {
    if (stg_cnt) {
      const uint4 b = (stg_k >> 2) ? b1 : b0;
      const uint32_t k = stg_k & 3u;
      const uint32_t d = k == 0u ? b.x : k == 1u ? b.y : k == 2u ? b.z : b.w;
      append(d, stg_skip, stg_cnt);
      stg_cnt = 0;
    }
  }
  // Stage the dword after the window (when it has room for it). Bytes come
  // from BLK_BYTES aligned blocks, one cache line: a lane requests each line
  // of its blob once, so the line cannot be evicted from the L2 (which cannot
  // hold every lane's current line) between two requests of it and fetched
  // again (with 16-byte blocks a line was fetched about 7 times). A block
  // holding a stream byte lies in the buffer's pages, like the dword did.
  __device__ __forceinline__ void refill() {
    if (stg_cnt == 0u && wn <= 4u) {
      const uint32_t q = qn;
      if (q >= n) {
        qn += 8u - wn;
        wn = 8;  // the rest of the window is the stream's zero tail
      } else {
        const uintptr_t a = (uintptr_t)(p + q);
        stg_skip = (uint32_t)(a & 3u);
        const uint32_t left = n - q;
        stg_cnt = 4u - stg_skip < left ? 4u - stg_skip : left;
        stg_k = (uint32_t)(a >> 2) & (BLK_BYTES / 4u - 1u);
        const uintptr_t ba = a & ~(uintptr_t)(BLK_BYTES - 1u);
        if (ba != blk_a) {
          const uint4 *bp = (const uint4 *)ba;
          b0 = bp[0], b1 = bp[1];
          blk_a = ba;
        }
      }
    }
  }
  __device__ __forceinline__ void fill_sync(uint32_t need) {
    merge();
    while (wn < need) {
      const uint32_t q = qn;
      if (q >= n) {
        qn += 8u - wn;
        wn = 8;
        break;
      }
      uint32_t sk, cn;
      const uint32_t d = load_at(q, sk, cn);
      append(d, sk, cn);
    }
  }
  // Shift k <= 2 bytes from the window into code.
  __device__ __forceinline__ void take(uint32_t k) {
    if (wn < k) fill_sync(k);
    code = (uint32_t)((((uint64_t)code << 32) | (w >> 32)) >> (32u - 8u * k));
    w <<= 8u * k;
    wn -= k;
  }
  __device__ __forceinline__ void init(const uint8_t *b, uint32_t len) {
    p = b;
    n = len;
    qn = 0;
    w = 0;
    wn = 0;
    stg_cnt = 0;
    blk_a = 0;
    code = 0;
    take(2);
    take(2);
    take(1);  // the first byte is shifted out of code, as in dct3d.c:227
    range = 0xFFFFFFFFu;
  }
  // Renormalise: bytes until range >= 2^24. range >= 2^8 on entry (a context
  // decision leaves at least range >> 12 >= 2^12 since its probability lies in
  // [1, 4095], a bypass 2^23, a raw-bits step of s <= 16 bits 2^8), so at most
  // two bytes, and their count is the leading-zero count / 8 (0 for
  // [2^24, 2^32), 1 for [2^16, 2^24), 2 for [2^8, 2^16)).
  __device__ __forceinline__ void norm() {
    const uint32_t k = (uint32_t)__clz((int)range) >> 3;
    take(k);
    range <<= 8u * k;
  }
  __device__ __forceinline__ uint32_t bit(uint32_t ci) {
    uint16_t *pc = ctx + ci * BLOB_NT;
    const uint32_t p0 = *pc;
    const uint32_t r0 = (range >> 12) * p0;
    uint32_t b;
    if (code < r0) {
      range = r0;
      *pc = (uint16_t)(p0 + ((4096u - p0) >> 4));
      b = 0;
    } else {
      code -= r0;
      range -= r0;
      *pc = (uint16_t)(p0 - (p0 >> 4));
      b = 1;
    }
    norm();
    return b;
  }
  // One decision of either kind with the same instruction stream, free of
  // branches: an adaptive context bit (ci) or, when byp, an equiprobable
  // bypass bit (ci is then any context, written back unchanged).
  __device__ __forceinline__ uint32_t decide(uint32_t ci, bool byp) {
    uint16_t *pc = ctx + ci * BLOB_NT;
    const uint32_t p0 = *pc;
    // every arm computed first, so each choice is a select, not a branch
    const uint32_t r_ctx = (range >> 12) * p0, r_byp = range >> 1;
    const uint32_t r0 = __builtin_unpredictable(byp) ? r_byp : r_ctx;
    const bool b = code >= r0;
    const uint32_t code1 = code - r0, range1 = range - r0;
    const uint32_t p_one = p0 - (p0 >> 4), p_zero = p0 + ((4096u - p0) >> 4);
    code = __builtin_unpredictable(b) ? code1 : code;
    range = __builtin_unpredictable(b && !byp) ? range1 : r0;
    const uint32_t up = __builtin_unpredictable(b) ? p_one : p_zero;
    *pc = (uint16_t)(__builtin_unpredictable(byp) ? p0 : up);
    norm();
    return b ? 1u : 0u;
  }
  __device__ __forceinline__ uint32_t bypass() {
    range >>= 1;
    const uint32_t b = code >= range ? 1u : 0u;
    if (b) code -= range;
    norm();
    return b;
  }
  __device__ __forceinline__ uint32_t bits(int k) {  // k raw bits, most significant first
    uint32_t v = 0;
    while (k > 0) {
      const int s = k > 16 ? 16 : k;
      range >>= s;
      uint32_t q = code / range;
      const uint32_t m = (1u << s) - 1u;
      if (q > m) q = m;
      code -= q * range;
      norm();
      v = (v << s) | q;
      k -= s;
    }
    return v;
  }
  __device__ __forceinline__ uint32_t eg() {
    uint32_t nb = 0;
    while (bypass())
      if (++nb > 31u) return 0;
    return nb ? ((1u << nb) | bits((int)nb)) - 1u : 0u;
  }
  __device__ __forceinline__ int32_t seg() {
    const uint32_t z = eg();
    return (z & 1u) ? -(int32_t)((z + 1u) >> 1) : (int32_t)(z >> 1);
  }
  __device__ __forceinline__ float fval() {
    if (!bypass()) return (float)seg();
    return __uint_as_float(bits(32));
  }
};

struct ParseArgs {
  uint32_t n, dtype, cap;
  const uint8_t *bytes;
  const uint64_t *offsets;
  const uint32_t *lengths;
  r3d_chunk_header *hdr;
  float *steps;
  uint32_t *pair_off, *pair_cnt, *pairs;
  uint32_t *corr_off, *corr_cnt;
  uint16_t *corr_pos;
  float *corr_delta;
  int32_t *status;
  const uint32_t *order;  // lane -> blob (NULL: lane i parses blob i)
};

__global__ __launch_bounds__(BLOB_NT) void k_parse_blobs(ParseArgs a) {
  __shared__ uint16_t s_scan[4096];
  __shared__ __attribute__((aligned(16))) uint16_t s_ctx[NCTX * BLOB_NT];
  const uint32_t t = threadIdx.x;
  for (uint32_t i = t; i < 4096u; i += BLOB_NT) s_scan[i] = c_scan[i];
  __syncthreads();
  const uint32_t lane = blockIdx.x * BLOB_NT + t;
  // lanes past the batch stay to the workgroup barrier below, doing nothing
  const bool in = lane < a.n;
  const uint32_t c = in ? (a.order ? a.order[lane] : lane) : 0u;
  RDec d;
  d.ctx = s_ctx + t;
  const uint64_t base = (uint64_t)c * a.cap;
  const uint8_t *b = in ? a.bytes + a.offsets[c] : nullptr;
  const uint32_t len = in ? a.lengths[c] : 0u;
  int32_t st = 1;
  uint32_t np = 0, nc = 0;
  float vmin = 0.0f, vspan = 0.0f, qstep = 0.0f;
  int32_t dcq = 0;
  uint32_t has_corr = 0, eob = 0;
  if (!in || len < 2u || b[0] != 0xD3u || b[1] != a.dtype) {
    st = 0;
  } else {
    d.init(b + 2, len - 2u);
    vmin = d.fval();
    vspan = d.fval();
    qstep = d.fval();
    dcq = d.seg();
    const float dc = (float)dcq / 16.0f;
    if (!(vspan > 0.0f) || !(qstep > 0.0f) || !__builtin_isfinite(vmin) || !__builtin_isfinite(dc)) st = 0;
  }
  if (st == 1) {
    for (uint32_t i = 0; i < NCTX; i++) d.ctx[i * BLOB_NT] = c_prior[i];
    has_corr = d.bit(CTX_FLAG);
    // EOB (dct3d.c:365-371): up to 13 context bits, then k - 1 raw bits
    uint32_t k = 0;
    while (k < 13u && d.bit(CTX_EOB + k)) k++;
    eob = k <= 1u ? k : ((1u << (k - 1u)) | d.bits((int)k - 1));
    if (eob > 4096u) eob = 4096u;
  }
  // Every lane of the workgroup is past its EOB from here on, and nothing
  // after reads the EOB contexts again: their rows become the pair staging
  // area, in slots that cover other lanes' halfwords.
  __syncthreads();
  if (st == 1) {
    // Levels (dct3d.c:390-405) as a per-lane state machine that makes exactly
    // one binary decision per iteration: significance in context (band,
    // min(3, popcount of the last 16 decisions)), the unary magnitude ladder
    // (contexts 0-10, escape flag 11), the Exp-Golomb escape prefix (bypass),
    // the sign (bypass). Lanes of a wave stay converged whatever each blob's
    // mix of zero runs and ladders; the last position's significance is
    // implied.
    // The window's refills are issued every 8 iterations and merged 8
    // iterations later.
    enum : uint32_t { PH_SIG, PH_MAG, PH_EGP, PH_SIGN, PH_DONE };
    uint32_t p = 0, recent = 0, m = 0, nb = 0;
    // Pairs go out 8 at a time (two 16-byte stores filling a 32-byte
    // sector) when the capacity and the array keep such groups aligned: a
    // scattered 4-byte store per pair made the memory system fetch and write
    // whole sectors, and on this ISA every store also holds up the next wait
    // for a window refill. A lane stages its pairs 0..6 of a group in LDS, in
    // the EOB context rows (44..57: 14 x 256 halfwords = 7 x 256 dwords),
    // free for every lane since the barrier above (a slot covers the
    // halfwords of two other lanes, which may sit in another wave: without
    // the barrier a wave that was still at its priors or its EOB turned a
    // staged -1 level into +1); the eighth pair completes the group from a
    // register. (A register shift of 8 costs 8 moves in nearly every
    // iteration, since some lane of the wave is nearly always at a sign;
    // groups of 4 in registers wrote each sector twice.)
    const bool vec8 = (a.cap & 7u) == 0u && ((uintptr_t)a.pairs & 31u) == 0u;
    uint32_t *stg = (uint32_t *)(s_ctx + CTX_EOB * BLOB_NT) + t;  // slot j at stg[j * BLOB_NT]
    const auto stg_put = [&](uint32_t j, uint32_t v) { stg[j * BLOB_NT] = v; };
    const auto stg_get = [&](uint32_t j) -> uint32_t { return stg[j * BLOB_NT]; };
    uint32_t e = s_scan[0];
    uint32_t ph = eob == 0u ? PH_DONE : eob == 1u ? PH_MAG : PH_SIG;
    if (eob == 1u) recent = 1u;
    d.refill();
    for (uint32_t it = 1; ph != PH_DONE; it++) {
      uint32_t dens = __builtin_popcount(recent & 0xFFFFu);
      dens = dens < 3u ? dens : 3u;
      const bool sig = ph == PH_SIG, mag = ph == PH_MAG, egp = ph == PH_EGP, sgn = ph == PH_SIGN;
      const uint32_t ci_sig = CTX_SIG + (e >> 12) * 4u + dens, ci_mag = CTX_MAG + m;
      const uint32_t ci_oth = __builtin_unpredictable(mag) ? ci_mag : (uint32_t)CTX_FLAG;
      const uint32_t ci = __builtin_unpredictable(sig) ? ci_sig : ci_oth;
      const uint32_t b = d.decide(ci, egp || sgn);
      if (egp) {
        nb += b;
        if (!b || nb > 31u) {
          // nb can reach 32 on a corrupt stream; the shift count is taken mod
          // 32, as the x86 and gfx9 shift instructions do for dct3d.c:326
          m = 12u + (nb ? ((1u << (nb & 31u)) | d.bits((int)nb)) - 1u : 0u);
          ph = PH_SIGN;
        }
      }
      if (sgn) {
        if (m > (b ? 32768u : 32767u)) st = -1;  // beyond the int16 level formats
        if (np >= a.cap) {
          st = st == 1 ? -2 : st;  // more nonzero levels than the capacity
        } else {
          const uint32_t lv = (uint32_t)(uint16_t)(int16_t)(b ? -(int32_t)m : (int32_t)m);
          const uint32_t pr = (e & 4095u) | (lv << 16);
          if (vec8) {
            const uint32_t slot = np & 7u;
            if (slot == 7u) {
              uint4 *q = (uint4 *)(a.pairs + base + np - 7u);
              q[0] = make_uint4(stg_get(0), stg_get(1), stg_get(2), stg_get(3));
              q[1] = make_uint4(stg_get(4), stg_get(5), stg_get(6), pr);
            } else {
              stg_put(slot, pr);
            }
          } else {
            a.pairs[base + np] = pr;
          }
          np++;
        }
      }
      // next state: SIG -> MAG (1) or next position (0); MAG -> SIGN with
      // magnitude m + 1 (0), m + 1 (1, m < 11) or EGP (1, m == 11); SIGN ->
      // next position. The last position's significance is implied.
      const bool adv = (sig && !b) || sgn;
      const uint32_t pn = p + (adv ? 1u : 0u);
      const uint32_t en = s_scan[pn < 4095u ? pn : 4095u];
      recent = __builtin_unpredictable(sig) ? (recent << 1) | b : recent;
      const bool implied = adv && pn + 1u == eob;
      recent = __builtin_unpredictable(implied) ? (recent << 1) | 1u : recent;
      const uint32_t m1 = m + 1u;
      m = __builtin_unpredictable(sig || implied) ? 0u : __builtin_unpredictable(mag) ? m1 : m;
      nb = __builtin_unpredictable(mag) ? 0u : nb;
      const uint32_t from_mag = __builtin_unpredictable(!b) ? PH_SIGN : m1 < 12u ? PH_MAG : PH_EGP;
      const uint32_t at_adv = __builtin_unpredictable(pn >= eob) ? PH_DONE : implied ? PH_MAG : PH_SIG;
      const uint32_t at_stay = __builtin_unpredictable(sig) ? PH_MAG : __builtin_unpredictable(mag) ? from_mag : ph;
      ph = __builtin_unpredictable(adv) ? at_adv : at_stay;
      p = pn;
      e = en;
      if ((it & 7u) == 0u) {
        d.merge();
        d.refill();
      }
    }
    if (vec8) {  // the last np % 8 pairs are staged in slots 0 .. np % 8 - 1
      const uint32_t r = np & 7u;
      for (uint32_t k = 0; k < r; k++) a.pairs[base + np - r + k] = stg_get(k);
    }
    if (has_corr) {  // dct3d.c:572-587
      uint32_t cnt = d.eg() + 1u, pos = 0;
      if (cnt > 4096u) cnt = 4096u;
      float cq = d.fval();
      if (!(cq > 0.0f) || !__builtin_isfinite(cq)) cq = 1.0f;
      for (uint32_t i = 0; i < cnt; i++) {
        pos += d.eg();
        const uint32_t neg = d.bypass();
        const uint32_t mag = d.eg() + 1u;
        if (pos >= 4096u) break;
        const float delta = (float)mag * cq;
        if (nc >= a.cap) {
          st = st == 1 ? -2 : st;
          break;
        }
        a.corr_pos[base + nc] = (uint16_t)pos;
        a.corr_delta[base + nc] = neg ? -delta : delta;
        nc++;
      }
    }
  }
  if (!in) return;
  r3d_chunk_header h;
  h.vmin = vmin;
  h.vspan = st == 1 ? vspan : 0.0f;  // vspan 0 makes the decoder zero-fill the chunk (status 0)
  h.dcq = dcq;
  h.table = c;
  a.hdr[c] = h;
  for (uint32_t f = 0; f < 46u; f++) a.steps[(size_t)c * 46u + f] = st == 1 ? qstep * c_hfw[f] : 0.0f;
  a.pair_off[c] = (uint32_t)base;
  a.corr_off[c] = (uint32_t)base;
  a.pair_cnt[c] = st == 1 ? np : 0u;
  a.corr_cnt[c] = st == 1 ? nc : 0u;
  if (c + 1u == a.n) {
    a.pair_off[a.n] = (uint32_t)((uint64_t)a.n * a.cap);
    a.corr_off[a.n] = (uint32_t)((uint64_t)a.n * a.cap);
  }
  if (a.status) a.status[c] = st;
}

// ---- encode side (workloads) -------------------------------------------------
struct REnc {
  uint8_t *out;
  uint64_t cap, n;
  uint64_t low;
  uint32_t range;
  uint32_t cache;
  uint64_t pending;
  uint16_t *ctx;

  __device__ __forceinline__ void put(uint32_t b) {
    if (n < cap) out[n] = (uint8_t)b;
    n++;
  }
  __device__ __forceinline__ void shift() {  // dct3d.c:188-195
    if ((uint32_t)(low >> 32) != 0u || low < 0xFF000000ull) {
      const uint32_t carry = (uint32_t)(low >> 32);
      do {
        put((cache + carry) & 0xFFu);
        cache = 0xFFu;
      } while (--pending);
      cache = (uint32_t)(low >> 24) & 0xFFu;
    }
    pending++;
    low = (low << 8) & 0xFFFFFFFFull;
  }
  __device__ __forceinline__ void norm() {
    while (range < RC_TOP) {
      shift();
      range <<= 8;
    }
  }
  __device__ __forceinline__ void bit(uint32_t ci, uint32_t b) {
    uint16_t *pc = ctx + ci * BLOB_NT;
    const uint32_t p0 = *pc;
    const uint32_t r0 = (range >> 12) * p0;
    if (!b) {
      range = r0;
      *pc = (uint16_t)(p0 + ((4096u - p0) >> 4));
    } else {
      low += r0;
      range -= r0;
      *pc = (uint16_t)(p0 - (p0 >> 4));
    }
    norm();
  }
  __device__ __forceinline__ void bypass(uint32_t b) {
    range >>= 1;
    if (b) low += range;
    norm();
  }
  __device__ __forceinline__ void bits(uint32_t v, int k) {
    while (k > 0) {
      const int s = k > 16 ? 16 : k;
      range >>= s;
      low += (uint64_t)((v >> (k - s)) & ((1u << s) - 1u)) * range;
      norm();
      k -= s;
    }
  }
  __device__ __forceinline__ void eg(uint32_t v) {
    uint32_t nb = 0, t = v + 1u;
    while (t > 1u) {
      t >>= 1;
      nb++;
    }
    for (uint32_t i = 0; i < nb; i++) bypass(1);
    bypass(0);
    if (nb) bits((v + 1u) & ((1u << nb) - 1u), (int)nb);
  }
  __device__ __forceinline__ void seg(int32_t v) { eg(v < 0 ? (uint32_t)(-2 * v - 1) : (uint32_t)(2 * v)); }
  __device__ __forceinline__ void fval(float f) {  // dct3d.c:266-271
    const float r = __builtin_truncf(f);
    if (r == f && __builtin_fabsf(f) < 16777216.0f) {
      bypass(0);
      seg((int32_t)r);
    } else {
      bypass(1);
      bits(__float_as_uint(f), 32);
    }
  }
};

struct EncodeBlobArgs {
  uint32_t n, dtype;
  const int16_t *levels;
  const r3d_chunk_header *hdr;
  const float *qsteps;
  uint8_t *out;
  uint64_t slot;
  uint32_t *lengths;
};

__global__ __launch_bounds__(BLOB_NT) void k_encode_blobs(EncodeBlobArgs a) {
  __shared__ uint16_t s_scan[4096];
  __shared__ uint16_t s_ctx[NCTX * BLOB_NT];
  const uint32_t t = threadIdx.x;
  for (uint32_t i = t; i < 4096u; i += BLOB_NT) s_scan[i] = c_scan[i];
  __syncthreads();
  const uint32_t c = blockIdx.x * BLOB_NT + t;
  if (c >= a.n) return;
  const int16_t *lv = a.levels + (size_t)c * 4096u;
  uint8_t *o = a.out + (uint64_t)c * a.slot;
  if (a.slot >= 1u) o[0] = 0xD3u;
  if (a.slot >= 2u) o[1] = (uint8_t)a.dtype;
  REnc e;
  e.out = o + 2;
  e.cap = a.slot >= 2u ? a.slot - 2u : 0u;
  e.n = 0;
  e.low = 0;
  e.range = 0xFFFFFFFFu;
  e.cache = 0;
  e.pending = 1;
  e.ctx = s_ctx + t;
  for (uint32_t i = 0; i < NCTX; i++) e.ctx[i * BLOB_NT] = c_prior[i];
  const r3d_chunk_header h = a.hdr[c];
  e.fval(h.vmin);
  e.fval(h.vspan);
  e.fval(a.qsteps[c]);
  e.seg(h.dcq);
  e.bit(CTX_FLAG, 0);  // no corrections
  uint32_t eob = 0;
  for (uint32_t p = 4096u; p-- > 0;)
    if (lv[s_scan[p] & 4095u] != 0) {
      eob = p + 1u;
      break;
    }
  uint32_t k = 0;  // dct3d.c:357-363
  while ((1u << k) <= eob) k++;
  for (uint32_t i = 0; i < k; i++) e.bit(CTX_EOB + i, 1);
  if (k < 13u) e.bit(CTX_EOB + k, 0);
  if (k > 1u) e.bits(eob & ((1u << (k - 1u)) - 1u), (int)k - 1);
  uint32_t recent = 0;
  for (uint32_t p = 0; p < eob; p++) {
    const uint32_t se = s_scan[p];
    const int32_t v = lv[se & 4095u];
    uint32_t dens = __builtin_popcount(recent & 0xFFFFu);
    dens = dens < 3u ? dens : 3u;
    if (p + 1u != eob) e.bit(CTX_SIG + (se >> 12) * 4u + dens, v != 0);
    recent = (recent << 1) | (v != 0 ? 1u : 0u);
    if (!v) continue;
    uint32_t rest = (uint32_t)(v < 0 ? -v : v) - 1u, j = 0;  // dct3d.c:310-319
    while (j < 11u && rest > 0u) {
      e.bit(CTX_MAG + j, 1);
      rest--;
      j++;
    }
    if (rest == 0u) {
      e.bit(CTX_MAG + j, 0);
    } else {
      e.bit(CTX_MAG + 11u, 1);
      e.eg(rest);
    }
    e.bypass(v < 0 ? 1u : 0u);
  }
  for (int i = 0; i < 5; i++) e.shift();
  a.lengths[c] = (uint32_t)(e.n + 2u);
}

int upload_blob_tables(r3d_hip_ctx *ctx) {
  uint16_t scan[4096];
  int start[47] = {0};
  for (int i = 0; i < 4096; i++) start[(i >> 8) + ((i >> 4) & 15) + (i & 15) + 1]++;
  for (int f = 1; f < 47; f++) start[f] += start[f - 1];
  for (int i = 0; i < 4096; i++) {
    const int f = (i >> 8) + ((i >> 4) & 15) + (i & 15);
    const int band = f * 8 / 48;  // dct3d.c:352-355
    scan[start[f]++] = (uint16_t)(i | (band > 7 ? 7 : band) << 12);
  }
  float hfw[46];
  for (int f = 0; f < 46; f++) hfw[f] = powf(1.0f + (float)f, 0.65f);  // the host libm, as the CPU decoder
  R3D_HIP_CHECK(ctx, hipMemcpyToSymbol(HIP_SYMBOL(c_scan), scan, sizeof scan, 0, hipMemcpyHostToDevice));
  R3D_HIP_CHECK(ctx, hipMemcpyToSymbol(HIP_SYMBOL(c_hfw), hfw, sizeof hfw, 0, hipMemcpyHostToDevice));
  return R3D_OK;
}

__global__ __launch_bounds__(256) void k_iota(uint32_t *v, uint32_t n) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i < n) v[i] = i;
}

}  // namespace r3d

using namespace r3d;

// The context's scratch for the parse's lane order. A larger batch takes a
// new buffer (at least twice the old one) and the old one is retired, not
// freed, until r3d_hip_destroy: a HIP graph that captured an earlier parse
// keeps its address and stays valid. A capture cannot allocate, so growing
// is refused while the stream is capturing.
static int sort_scratch(r3d_hip_ctx *ctx, size_t bytes, hipStream_t st, void **out) {
  if (ctx->sort_cap < bytes) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    R3D_HIP_CHECK(ctx, hipStreamIsCapturing(st, &cs));
    if (cs != hipStreamCaptureStatusNone)
      return r3d_fail(ctx, R3D_ERR_INVALID,
                      "parse_blobs: the lane-order scratch would grow to %zu bytes while the stream is capturing; "
                      "parse a batch this large once before capturing", bytes);
    const uint32_t nret = (uint32_t)(sizeof ctx->retired / sizeof ctx->retired[0]);
    if (ctx->d_sort && ctx->nretired == nret)
      return r3d_fail(ctx, R3D_ERR_NOMEM, "parse_blobs: %u retired order buffers", nret);
    const size_t cap = bytes > 2u * ctx->sort_cap ? bytes : 2u * ctx->sort_cap;
    void *p = nullptr;
    if (hipMalloc(&p, cap) != hipSuccess)
      return r3d_fail(ctx, R3D_ERR_NOMEM, "parse_blobs: order scratch of %zu bytes", cap);
    if (ctx->d_sort) ctx->retired[ctx->nretired++] = ctx->d_sort;
    ctx->d_sort = p;
    ctx->sort_cap = cap;
  }
  *out = ctx->d_sort;
  return R3D_OK;
}

// ---- the allocation-free parse ----------------------------------------------
// r3d_hip_parse_blobs_noalloc keeps the lane order and the radix sort's
// temporary storage in this file-scope device array instead of context
// memory: one copy per GPU, resident from module load until the process
// ends. In host code `scratch` names a host-side stand-in, never the device
// memory: the device address comes from hipGetSymbolAddress, taken once per
// device (parse_noalloc_setup, from r3d_hip_init), and only that address is
// handed to kernels and to the sort. Nothing in it needs to start at zero:
// the iota and the sort write every word they read.
constexpr size_t NOALLOC_SCRATCH_BYTES = 16u << 20;
static __device__ __align__(256) uint8_t scratch[NOALLOC_SCRATCH_BYTES];

constexpr int NOALLOC_MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[NOALLOC_MAX_DEVICES];  // per device: the address of `scratch`

int r3d::parse_noalloc_setup(r3d_hip_ctx *ctx) {
  if (ctx->device >= NOALLOC_MAX_DEVICES)
    return r3d_fail(ctx, R3D_ERR_RANGE, "device %d beyond the %d the parse scratch table holds", ctx->device,
                    NOALLOC_MAX_DEVICES);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[ctx->device]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return r3d_fail(ctx, R3D_ERR_DEVICE, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    g_scratch_ptr[ctx->device] = p;
  }
  ctx->d_noalloc = g_scratch_ptr[ctx->device];
  return R3D_OK;
}

// r3d_hip_parse_blobs and r3d_hip_parse_blobs_noalloc: the lane order comes
// from the context's scratch (grown on demand) or from the fixed array.
static int parse_impl(r3d_hip_ctx *ctx, const r3d_blob_batch *in, const r3d_blob_parse_out *out, void *stream,
                      bool noalloc) {
  const char *fn = noalloc ? "parse_blobs_noalloc" : "parse_blobs";
  if (!ctx) return R3D_ERR_INVALID;
  if (!in || !out) return r3d_fail(ctx, R3D_ERR_INVALID, "%s: NULL batch or output", fn);
  if (in->nchunks == 0) return R3D_OK;
  if (!in->bytes || !in->offsets || !in->lengths)
    return r3d_fail(ctx, R3D_ERR_INVALID, "%s: bytes, offsets and lengths are required", fn);
  if (in->max_pairs == 0 || in->max_pairs > 4096u)
    return r3d_fail(ctx, R3D_ERR_INVALID, "%s: max_pairs %u outside 1..4096", fn, in->max_pairs);
  if ((uint64_t)in->nchunks * in->max_pairs > 0xFFFFFFFFull)
    return r3d_fail(ctx, R3D_ERR_RANGE, "%s: nchunks * max_pairs exceeds 32-bit offsets", fn);
  if (noalloc && in->nchunks > R3D_PARSE_NOALLOC_MAX_BLOBS)
    return r3d_fail(ctx, R3D_ERR_RANGE, "%s: %u blobs, above the %u its fixed scratch holds", fn, in->nchunks,
                    R3D_PARSE_NOALLOC_MAX_BLOBS);
  if (!out->headers || !out->step_tables || !out->pair_offsets || !out->pair_counts || !out->pairs ||
      !out->corr_offsets || !out->corr_counts || !out->corr_pos || !out->corr_delta)
    return r3d_fail(ctx, R3D_ERR_INVALID, "%s: every output array is required", fn);
  if (noalloc && !ctx->d_noalloc) return r3d_fail(ctx, R3D_ERR_DEVICE, "%s: no scratch address", fn);
  R3D_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  hipStream_t st = (hipStream_t)stream;
  // Lanes take the blobs longest first (a stable radix sort of the lengths):
  // a wave runs as long as its longest blob, and in chunk order a wave's
  // longest blob was 1.29x its mean (at 1024^3, qstep 8), after the sort
  // 1.0003x. The outputs stay indexed by blob, so any order is correct: the
  // sort uses the low 16 bits (two radix passes; a 3ddct blob is at most
  // 33,024 bytes, and longer ones are rejected wherever they land).
  const uint32_t n = in->nchunks;
  const uint32_t *order = nullptr;
  if (n >= 2u * BLOB_NT && n <= 0x7FFFFFFFu) {
    size_t temp = 0;
    R3D_HIP_CHECK(ctx, hipcub::DeviceRadixSort::SortPairsDescending(nullptr, temp, in->lengths, (uint32_t *)nullptr,
                                                                    (const uint32_t *)nullptr, (uint32_t *)nullptr,
                                                                    (int)n, 0, 16, st));
    const size_t arr = ((size_t)n * 4u + 255u) & ~(size_t)255u;
    const size_t need = 3u * arr + temp;
    void *ws = nullptr;
    if (noalloc) {
      if (need > NOALLOC_SCRATCH_BYTES)
        return r3d_fail(ctx, R3D_ERR_RANGE, "%s: the lane order of %u blobs needs %zu bytes, above the fixed %zu",
                        fn, n, need, NOALLOC_SCRATCH_BYTES);
      ws = ctx->d_noalloc;
    } else {
      const int rc = sort_scratch(ctx, need, st, &ws);
      if (rc) return rc;
    }
    uint8_t *p = (uint8_t *)ws;
    uint32_t *keys = (uint32_t *)p, *idx = (uint32_t *)(p + arr), *ord = (uint32_t *)(p + 2u * arr);
    hipLaunchKernelGGL(k_iota, dim3((n + 255u) / 256u), dim3(256), 0, st, idx, n);
    R3D_HIP_CHECK(ctx, hipcub::DeviceRadixSort::SortPairsDescending(p + 3u * arr, temp, in->lengths, keys,
                                                                    (const uint32_t *)idx, ord, (int)n, 0, 16, st));
    order = ord;
  }
  ParseArgs a{in->nchunks,       in->dtype,        in->max_pairs,     in->bytes,         in->offsets,
              in->lengths,       out->headers,     out->step_tables,  out->pair_offsets, out->pair_counts,
              out->pairs,        out->corr_offsets, out->corr_counts, out->corr_pos,     out->corr_delta,
              out->status,       order};
  hipLaunchKernelGGL(k_parse_blobs, dim3((in->nchunks + BLOB_NT - 1) / BLOB_NT), dim3(BLOB_NT), 0, st, a);
  R3D_HIP_CHECK(ctx, hipGetLastError());
  return R3D_OK;
}

extern "C" int r3d_hip_parse_blobs(r3d_hip_ctx *ctx, const r3d_blob_batch *in, const r3d_blob_parse_out *out,
                                   void *stream) {
  return parse_impl(ctx, in, out, stream, false);
}

extern "C" int r3d_hip_parse_blobs_noalloc(r3d_hip_ctx *ctx, const r3d_blob_batch *in,
                                           const r3d_blob_parse_out *out, void *stream) {
  return parse_impl(ctx, in, out, stream, true);
}

extern "C" int r3d_hip_encode_blobs(r3d_hip_ctx *ctx, uint32_t nchunks, const int16_t *levels,
                                    const r3d_chunk_header *headers, const float *qsteps, uint32_t dtype,
                                    uint8_t *out, uint64_t slot_bytes, uint32_t *lengths, void *stream) {
  if (!ctx) return R3D_ERR_INVALID;
  if (nchunks == 0) return R3D_OK;
  if (!levels || !headers || !qsteps || !out || !lengths)
    return r3d_fail(ctx, R3D_ERR_INVALID, "encode_blobs: NULL argument");
  if (slot_bytes < 2u) return r3d_fail(ctx, R3D_ERR_INVALID, "encode_blobs: slot_bytes below the 2-byte preamble");
  if (dtype == 0u || dtype > 255u) return r3d_fail(ctx, R3D_ERR_INVALID, "encode_blobs: dtype %u", dtype);
  R3D_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  EncodeBlobArgs a{nchunks, dtype, levels, headers, qsteps, out, slot_bytes, lengths};
  hipLaunchKernelGGL(k_encode_blobs, dim3((nchunks + BLOB_NT - 1) / BLOB_NT), dim3(BLOB_NT), 0,
                     (hipStream_t)stream, a);
  R3D_HIP_CHECK(ctx, hipGetLastError());
  return R3D_OK;
}
