"""ctypes access to the CPU oracle (liboracle.so). TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, as the checker / the timed CPU baseline. The product never does.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

f32p = np.ctypeslib.ndpointer(np.float32, flags="C")
u8p = np.ctypeslib.ndpointer(np.uint8, flags="C")
i16p = np.ctypeslib.ndpointer(np.int16, flags="C")
u16p = np.ctypeslib.ndpointer(np.uint16, flags="C")
u32p = np.ctypeslib.ndpointer(np.uint32, flags="C")
i32p = np.ctypeslib.ndpointer(np.int32, flags="C")
VP = C.c_void_p


class Counters(C.Structure):
    _fields_ = [("samples", C.c_uint64), ("skips", C.c_uint64), ("shaded", C.c_uint64), ("rays", C.c_uint64),
                ("blended", C.c_uint64), ("blended_shaded", C.c_uint64)]


class BricksC(C.Structure):
    _fields_ = [("page", C.c_void_p), ("npage", C.c_uint32), ("atlas", C.c_void_p), ("abpa", C.c_uint32)]


class VolumeC(C.Structure):
    _fields_ = [("nlev", C.c_uint32), ("dims", (C.c_uint32 * 3) * 16), ("lev", C.c_void_p * 16),
                ("odim", C.c_uint32 * 3), ("occ", C.c_void_p), ("touched", C.c_void_p)]


_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} missing: run `make -C oracle`")
    L = C.CDLL(LIB_PATH)
    L.r3do_dct_basis.argtypes = [f32p]
    L.r3do_step_table_qstep.argtypes = [C.c_float, f32p]
    L.r3do_step_table_linear.argtypes = [f32p]
    L.r3do_decode_chunk.restype = C.c_int
    L.r3do_decode_chunk.argtypes = [i16p, VP, f32p, VP, VP, C.c_uint32, u8p, VP]
    L.r3do_dequant.argtypes = [i16p, f32p, f32p]
    L.r3do_idct3.argtypes = [f32p, f32p]
    L.r3do_fdct3.argtypes = [f32p, f32p]
    L.r3do_decode_batch.argtypes = [C.c_uint32, i16p, VP, f32p, C.c_uint32, VP, VP, VP, u8p, VP, C.c_int]
    L.r3do_encode_chunk.argtypes = [u8p, f32p, i16p, VP]
    L.r3do_encode_volume.argtypes = [u8p, C.c_uint32, C.c_uint32, C.c_uint32, f32p, i16p, VP, C.c_int]
    L.r3do_downsample.argtypes = [u8p, C.c_uint32, C.c_uint32, C.c_uint32, u8p]
    L.r3do_occupancy.argtypes = [u8p, C.c_uint32, C.c_uint32, C.c_uint32, u8p]
    L.r3do_raycast.restype = C.c_int
    L.r3do_raycast.argtypes = [C.POINTER(VolumeC), f32p, VP, C.c_uint32, C.c_float, f32p, C.c_uint32,
                               C.c_uint32, VP, C.c_uint32, C.c_uint32, C.c_int, C.POINTER(Counters), C.c_int]
    L.r3do_raycast_bricks.restype = C.c_int
    L.r3do_raycast_bricks.argtypes = [C.POINTER(BricksC), f32p, VP, C.c_uint32, C.c_float, f32p, C.c_uint32,
                                      C.c_uint32, VP, C.c_uint32, C.c_uint32, C.c_int, C.POINTER(Counters),
                                      C.c_int]
    L.r3do_plant.argtypes = [C.c_int]
    L.r3do_time_decoder.restype = C.c_double
    L.r3do_time_decoder.argtypes = [VP, u8p, VP, VP, C.c_uint32, C.c_uint32, C.c_int, C.POINTER(C.c_uint64),
                                    C.POINTER(C.c_uint64)]
    L.r3do_bricks_build.restype = C.c_uint32
    L.r3do_bricks_build.argtypes = [C.POINTER(VolumeC), VP, u32p, C.c_uint32, u8p, C.c_uint32]
    for fn in ("r3do_log2f", "r3do_exp2f"):
        getattr(L, fn).restype = C.c_float
        getattr(L, fn).argtypes = [C.c_float]
    L.r3do_pow_unit.restype = C.c_float
    L.r3do_pow_unit.argtypes = [C.c_float, C.c_float]
    L.r3do_hash12.restype = C.c_float
    L.r3do_hash12.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32]
    L.r3do_math_batch.argtypes = [f32p, C.c_uint32, C.c_uint32, f32p]
    f64p = np.ctypeslib.ndpointer(np.float64, flags="C")
    L.r3do_raycast_diff.restype = C.c_int
    L.r3do_raycast_diff.argtypes = [C.POINTER(VolumeC), f64p, f64p, u32p, VP, C.c_uint32, C.c_double, f64p,
                                    C.c_uint32, C.c_uint32, VP, VP, VP]
    L.r3do_raycast_diff_window.restype = C.c_int
    L.r3do_raycast_diff_window.argtypes = L.r3do_raycast_diff.argtypes + [C.c_uint32] * 4
    L.r3do_synth_sheets.argtypes = [u8p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint64,
                                    C.c_uint32, C.c_uint32, C.c_int]
    L.r3do_parse_blob.restype = C.c_int
    L.r3do_parse_blob.argtypes = [u8p, C.c_size_t, C.c_uint32, C.POINTER(BlobHdr), i16p, C.POINTER(C.c_uint32),
                                  np.ctypeslib.ndpointer(np.uint16, flags="C"), f32p]
    L.r3do_encode_blob.restype = C.c_size_t
    L.r3do_encode_blob.argtypes = [C.POINTER(BlobHdr), i16p, C.c_uint32, u8p, C.c_size_t]
    L.r3do_decode_blobs.argtypes = [C.c_uint32, u8p, VP, VP, C.c_uint32, u8p, VP, C.c_int]
    L.r3do_crc32c.restype = C.c_uint32
    L.r3do_crc32c.argtypes = [u8p, C.c_size_t]
    L.r3do_shard_valid.restype = C.c_int
    L.r3do_shard_valid.argtypes = [u8p, C.c_uint64]
    L.r3do_deblock_chunk_u8.argtypes = [u8p, VP, C.c_float, C.c_float, u8p]
    L.r3do_deblock_u8.argtypes = [u8p, C.c_uint64, C.c_uint64, C.c_uint64, C.c_float, C.c_float]
    L.r3do_shard_decode_region.restype = C.c_int
    L.r3do_shard_decode_region.argtypes = [VP, C.c_uint32] + [C.c_uint64] * 6 + [C.c_uint32] * 3 + [u8p, VP, VP,
                                                                                                  C.c_int]
    _lib = L
    return L


def hdr_array(vmin, vspan, dcq, table=None):
    """r3do_chunk_hdr[n] as int32 [n,4] (floats bit-cast)."""
    n = np.size(vmin)
    h = np.zeros((n, 4), np.int32)
    h[:, 0] = np.asarray(vmin, np.float32).reshape(-1).view(np.int32)
    h[:, 1] = np.asarray(vspan, np.float32).reshape(-1).view(np.int32)
    h[:, 2] = np.asarray(dcq, np.int32).reshape(-1)
    if table is not None:
        h[:, 3] = np.asarray(table, np.uint32).reshape(-1).view(np.int32)
    return h


def basis():
    out = np.zeros(256, np.float32)
    load().r3do_dct_basis(out)
    return out


def step_linear():
    s = np.zeros(46, np.float32)
    load().r3do_step_table_linear(s)
    return s


def step_qstep(q):
    s = np.zeros(46, np.float32)
    load().r3do_step_table_qstep(q, s)
    return s


def decode_chunk(levels, vmin, vspan, dcq, step46, corr_pos=None, corr_delta=None):
    """Returns (ok, u8[4096], pre_round f32[4096])."""
    L = load()
    h = hdr_array([vmin], [vspan], [dcq])
    out = np.zeros(4096, np.uint8)
    pre = np.zeros(4096, np.float32)
    n = 0 if corr_pos is None else len(corr_pos)
    cp = np.ascontiguousarray(corr_pos if n else np.zeros(1), np.uint16)
    cd = np.ascontiguousarray(corr_delta if n else np.zeros(1), np.float32)
    ok = L.r3do_decode_chunk(np.ascontiguousarray(levels, np.int16), h.ctypes.data, np.ascontiguousarray(step46, np.float32),
                             cp.ctypes.data, cd.ctypes.data, n, out, pre.ctypes.data)
    return ok, out, pre


def decode_batch(levels, hdr, tables, corr_off=None, corr_pos=None, corr_delta=None, nthreads=8):
    L = load()
    levels = np.ascontiguousarray(levels, np.int16).reshape(-1)
    n = levels.size // 4096
    out = np.zeros(n * 4096, np.uint8)
    status = np.zeros(n, np.int32)
    tables = np.ascontiguousarray(tables, np.float32).reshape(-1)
    hdr = np.ascontiguousarray(hdr, np.int32)
    args = [None, None, None]
    if corr_off is not None:
        co = np.ascontiguousarray(corr_off, np.uint32)
        cp = np.ascontiguousarray(corr_pos, np.uint16)
        cd = np.ascontiguousarray(corr_delta, np.float32)
        args = [co.ctypes.data, cp.ctypes.data, cd.ctypes.data]
    L.r3do_decode_batch(n, levels, hdr.ctypes.data, tables, tables.size // 46, args[0], args[1], args[2], out,
                        status.ctypes.data, nthreads)
    return out.reshape(n, 4096), status


def encode_volume(vol, nx, ny, nz, step46, nthreads=8):
    L = load()
    n = (nx // 16) * (ny // 16) * (nz // 16)
    lv = np.zeros(n * 4096, np.int16)
    hdr = np.zeros((n, 4), np.int32)
    L.r3do_encode_volume(np.ascontiguousarray(vol, np.uint8).reshape(-1), nx, ny, nz,
                         np.ascontiguousarray(step46, np.float32), lv, hdr.ctypes.data, nthreads)
    return lv.reshape(n, 4096), hdr


def synth_sheets(nx, ny, nz, nsheets, seed, nthreads=8, z0=0, z1=None):
    out = np.zeros(nx * ny * nz, np.uint8)
    load().r3do_synth_sheets(out, nx, ny, nz, nsheets, seed, z0, nz if z1 is None else z1, nthreads)
    return out


def pyramid(vol, nx, ny, nz, nlev):
    """Levels [(dims, array)] by the rounded 2x box."""
    L = load()
    levels = [((nx, ny, nz), np.ascontiguousarray(vol, np.uint8).reshape(-1))]
    for _ in range(1, nlev):
        (a, b, c), src = levels[-1]
        d = (max(a // 2, 1), max(b // 2, 1), max(c // 2, 1))
        dst = np.zeros(d[0] * d[1] * d[2], np.uint8)
        L.r3do_downsample(src, a, b, c, dst)
        levels.append((d, dst))
    return levels


def occupancy(vol, nx, ny, nz):
    o = ((nx + 7) // 8, (ny + 7) // 8, (nz + 7) // 8)
    out = np.zeros(o[0] * o[1] * o[2], np.uint8)
    load().r3do_occupancy(np.ascontiguousarray(vol, np.uint8).reshape(-1), nx, ny, nz, out)
    return o, out


def full_levels(nx, ny, nz):
    md = max(nx, ny, nz)
    n = 1
    while (md >> n) >= 1 and n < 16:
        n += 1
    return n


class OracleVolume:
    """Linear levels + occupancy, kept alive for the C struct's pointers."""

    def __init__(self, vol, nx, ny, nz, nlev=0, levels=None, occ=None):
        nlev = nlev or full_levels(nx, ny, nz)
        self.levels = levels if levels is not None else pyramid(vol, nx, ny, nz, nlev)
        self.odim, self.occ = occ if occ is not None else occupancy(vol, nx, ny, nz)
        s = VolumeC()
        s.nlev = len(self.levels)
        for l, (d, arr) in enumerate(self.levels):
            s.dims[l][0], s.dims[l][1], s.dims[l][2] = d
            s.lev[l] = arr.ctypes.data
        s.odim[0], s.odim[1], s.odim[2] = self.odim
        s.occ = self.occ.ctypes.data
        s.touched = None
        self.c = s
        self.touched = None

    def brick_count(self):
        """16^3 bricks over all levels (the bits of a touched-brick bitmap)."""
        return sum(-(-x // 16) * -(-y // 16) * -(-z // 16) for (x, y, z), _ in self.levels)

    def track_touched(self, on=True):
        """Start (or stop) marking the bricks the raycast's footprints read;
        returns the zeroed bitmap (uint32 words, level by level, x fastest)."""
        if on:
            self.touched = np.zeros((self.brick_count() + 31) // 32, np.uint32)
            self.c.touched = self.touched.ctypes.data
        else:
            self.touched = None
            self.c.touched = None
        return self.touched


QUALITY_FULL, QUALITY_FAST = 0, 1  # render_types.h:35-38

# planted mistakes (r3d_oracle.h R3DO_PLANT_*): sensitivity checks of the tests only
PLANT_NONE, PLANT_SPEC_EXP, PLANT_LIT_LERP, PLANT_TAP_RADIUS, PLANT_MIP_WEIGHT = 0, 1, 2, 3, 4
PLANT_SKIP_NUDGE, PLANT_SKIP_FACE, PLANT_FAST_RENORM, PLANT_KEY_LIGHT = 5, 6, 7, 8


class planted:
    """Context manager: the oracle with one planted mistake switched on."""

    def __init__(self, k):
        self.k = k

    def __enter__(self):
        load().r3do_plant(self.k)
        return self

    def __exit__(self, *exc):
        load().r3do_plant(PLANT_NONE)
        return False


def raycast(ovol, tf_f32, views, width, height, alpha_cutoff=0.98, row0=0, row_step=1, nthreads=8,
            px_steps=False, quality=QUALITY_FULL):
    """Returns (rgba float32 [h,w,4] (unrendered rows zero), counters dict, px_steps or None)."""
    L = load()
    n = len(views)
    arr = (type(views[0]) * n)(*views)
    img = np.zeros(height * width * 4, np.float32)
    steps = np.zeros(height * width, np.uint32) if px_steps else None
    cnt = Counters()
    rc = L.r3do_raycast(C.byref(ovol.c), np.ascontiguousarray(tf_f32, np.float32).reshape(-1), C.cast(arr, VP), n,
                        alpha_cutoff, img, width, height, None if steps is None else steps.ctypes.data, row0,
                        row_step, nthreads, C.byref(cnt), quality)
    if rc != 0:
        raise ValueError("oracle raycast: unsupported view")
    return (img.reshape(height, width, 4),
            {"samples": cnt.samples, "skips": cnt.skips, "shaded": cnt.shaded, "rays": cnt.rays,
             "blended": cnt.blended, "blended_shaded": cnt.blended_shaded},
            None if steps is None else steps.reshape(height, width))


BRICK = 128
SLOT_VOX = BRICK ** 3


def brick_page_words(level_dims):
    """Words of a page table: the 64-word header and one entry per 128^3 brick of every level."""
    return 64 + sum(-(-x // BRICK) * -(-y // BRICK) * -(-z // BRICK) for x, y, z in level_dims)


class OracleBricks:
    """A page table + slot atlas (r3do_bricks), kept alive for the C struct's pointers."""

    def __init__(self, page, atlas, abpa):
        self.page = np.ascontiguousarray(page, np.uint32)
        self.atlas = np.ascontiguousarray(atlas, np.uint8).reshape(-1)
        self.abpa = abpa
        assert self.atlas.size >= abpa ** 3 * SLOT_VOX
        self.c = BricksC(self.page.ctypes.data, self.page.size, self.atlas.ctypes.data, abpa)


def bricks_build(ovol, abpa, resident=None):
    """r3do_bricks_build over an OracleVolume's levels -> OracleBricks (slots
    taken in page order by the resident bricks)."""
    words = brick_page_words([d for d, _ in ovol.levels])
    page = np.zeros(words, np.uint32)
    atlas = np.zeros(abpa ** 3 * SLOT_VOX, np.uint8)
    res = None if resident is None else np.ascontiguousarray(resident, np.uint8)
    n = load().r3do_bricks_build(C.byref(ovol.c), None if res is None else res.ctypes.data, page, words, atlas,
                                 abpa ** 3)
    if n != words:
        raise ValueError("oracle bricks_build: more resident bricks than slots")
    return OracleBricks(page, atlas, abpa)


def raycast_bricks(ob, tf_f32, views, width, height, alpha_cutoff=0.98, row0=0, row_step=1, nthreads=8,
                   px_steps=False, quality=QUALITY_FULL):
    """r3do_raycast_bricks -> (rgba float32 [h,w,4], counters dict, px_steps or None)."""
    n = len(views)
    arr = (type(views[0]) * n)(*views)
    img = np.zeros(height * width * 4, np.float32)
    steps = np.zeros(height * width, np.uint32) if px_steps else None
    cnt = Counters()
    rc = load().r3do_raycast_bricks(C.byref(ob.c), np.ascontiguousarray(tf_f32, np.float32).reshape(-1),
                                    C.cast(arr, VP), n, alpha_cutoff, img, width, height,
                                    None if steps is None else steps.ctypes.data, row0, row_step, nthreads,
                                    C.byref(cnt), quality)
    if rc != 0:
        raise ValueError("oracle raycast_bricks: unsupported view")
    return (img.reshape(height, width, 4),
            {"samples": cnt.samples, "skips": cnt.skips, "shaded": cnt.shaded, "rays": cnt.rays},
            None if steps is None else steps.reshape(height, width))


def math_batch(x, fn):
    x = np.ascontiguousarray(x, np.float32)
    out = np.zeros_like(x)
    load().r3do_math_batch(x, x.size, fn, out)
    return out


class BlobHdr(C.Structure):
    _fields_ = [("vmin", C.c_float), ("vspan", C.c_float), ("qstep", C.c_float), ("dcq", C.c_int32)]


def parse_blob(blob, dtype=1):
    """r3do_parse_blob -> (status, (vmin, vspan, qstep, dcq), levels int16[4096],
    corr_pos uint16[n], corr_delta float32[n])."""
    blob = np.ascontiguousarray(blob, np.uint8)
    h = BlobHdr()
    lvl = np.zeros(4096, np.int16)
    nc = C.c_uint32()
    cp = np.zeros(4096, np.uint16)
    cd = np.zeros(4096, np.float32)
    st = load().r3do_parse_blob(blob if blob.size else np.zeros(1, np.uint8), blob.size, dtype, C.byref(h), lvl,
                                C.byref(nc), cp, cd)
    return st, (h.vmin, h.vspan, h.qstep, h.dcq), lvl, cp[:nc.value].copy(), cd[:nc.value].copy()


def encode_blob(vmin, vspan, qstep, dcq, levels, dtype=1, cap=4096 * 8 + 256):
    """r3do_encode_blob (no corrections) -> bytes as uint8 array."""
    h = BlobHdr(vmin, vspan, qstep, dcq)
    out = np.zeros(cap, np.uint8)
    n = load().r3do_encode_blob(C.byref(h), np.ascontiguousarray(levels, np.int16), dtype, out, cap)
    assert n <= cap
    return out[:n].copy()


def decode_blobs(data, offsets, lengths, dtype=1, nthreads=8):
    """r3do_decode_blobs: blobs (one byte array, per-blob offsets and lengths)
    -> (u8 [n, 4096], status [n])."""
    data = np.ascontiguousarray(data, np.uint8)
    offs = np.ascontiguousarray(offsets, np.uint64)
    lens = np.ascontiguousarray(lengths, np.uint32)
    n = offs.size
    out = np.zeros(n * 4096, np.uint8)
    status = np.zeros(n, np.int32)
    load().r3do_decode_blobs(n, data, offs.ctypes.data, lens.ctypes.data, dtype, out, status.ctypes.data, nthreads)
    return out.reshape(n, 4096), status


class ShardC(C.Structure):
    _fields_ = [("bytes", C.c_void_p), ("n", C.c_uint64), ("sz", C.c_uint32), ("sy", C.c_uint32),
                ("sx", C.c_uint32), ("reserved0", C.c_uint32)]


SHARD_GRID = 64
SHARD_INDEX_BYTES = SHARD_GRID ** 3 * 16 + 4


def crc32c(data):
    data = np.ascontiguousarray(data, np.uint8)
    return int(load().r3do_crc32c(data if data.size else np.zeros(1, np.uint8), data.size))


def shard_valid(shard):
    shard = np.ascontiguousarray(shard, np.uint8)
    return bool(load().r3do_shard_valid(shard if shard.size else np.zeros(1, np.uint8), shard.size))


def shard_decode_region(shards, dims, region, nthreads=8):
    """r3do_shard_decode_region. shards: list of (bytes uint8, sz, sy, sx);
    dims (nz, ny, nx); region (z0, y0, x0, dz, dy, dx). Returns (rc, dst
    uint8 [dz, dy, dx], counts [decoded, missing, corrupt, beyond int16],
    per-chunk status in the clamped chunk box)."""
    keep = [np.ascontiguousarray(b, np.uint8) for b, *_ in shards]
    arr = (ShardC * max(1, len(shards)))()
    for i, (b, (_, sz, sy, sx)) in enumerate(zip(keep, shards)):
        arr[i] = ShardC(b.ctypes.data if b.size else None, b.size, sz, sy, sx, 0)
    z0, y0, x0, dz, dy, dx = region
    nz, ny, nx = dims
    dst = np.zeros(int(dz) * int(dy) * int(dx), np.uint8)
    counts = np.zeros(4, np.uint32)
    nbox = 1
    if z0 < nz and y0 < ny and x0 < nx:
        for a0, d, n in ((z0, dz, nz), (y0, dy, ny), (x0, dx, nx)):
            a1 = min(a0 + d, n)
            nbox *= (a1 - 1) // 16 - a0 // 16 + 1
    status = np.zeros(nbox, np.int32)
    rc = load().r3do_shard_decode_region(C.addressof(arr), len(shards), nz, ny, nx, z0, y0, x0, dz, dy, dx, dst,
                                         counts.ctypes.data, status.ctypes.data, nthreads)
    return rc, dst.reshape(dz, dy, dx), counts, status


def shard_pack(blobs):
    """A shard in the layout of tools/3ddct/tools/export/shard.c:55-97: blobs
    (dict: chunk index (cz*64 + cy)*64 + cx -> uint8 array) back to back in
    chunk order, then the index ((~0, ~0) for a missing chunk) and its crc32c."""
    idx = np.full((SHARD_GRID ** 3, 2), np.iinfo(np.uint64).max, np.uint64)
    parts, off = [], 0
    for ci in sorted(blobs):
        b = np.ascontiguousarray(blobs[ci], np.uint8)
        idx[ci] = (off, b.size)
        parts.append(b)
        off += b.size
    ib = idx.astype("<u8").view(np.uint8).reshape(-1)
    crc = np.array([crc32c(ib)], "<u4").view(np.uint8)
    return np.concatenate(parts + [ib, crc]) if parts else np.concatenate([ib, crc])


def deblock(vol, step, strength=1.0):
    """r3do_deblock_u8 on a copy of vol (uint8 [nz, ny, nx])."""
    v = np.ascontiguousarray(vol, np.uint8).copy()
    nz, ny, nx = v.shape
    load().r3do_deblock_u8(v.reshape(-1), nz, ny, nx, step, strength)
    return v


def deblock_chunk(chunk, nbrs, step, strength=1.0):
    """r3do_deblock_chunk_u8: nbrs = [xlo, xhi, ylo, yhi, zlo, zhi], each a
    uint8[4096] or None."""
    keep = [None if b is None else np.ascontiguousarray(b, np.uint8).reshape(-1) for b in nbrs]
    arr = (C.c_void_p * 6)(*[None if b is None else b.ctypes.data for b in keep])
    out = np.zeros(4096, np.uint8)
    load().r3do_deblock_chunk_u8(np.ascontiguousarray(chunk, np.uint8).reshape(-1), C.cast(arr, C.c_void_p), step,
                                 strength, out)
    return out


def raycast_diff(ovol, tf, grid, gdims, views, width, height, alpha_cutoff=0.98, dimg=None, window=None):
    """FP64 differentiable forward (and reverse when dimg is given).
    window = (x0, y0, x1, y1) limits both to that ray window of each view.
    Returns (img [h,w,4] float64, dtf [256,4] or None, dgrid or None)."""
    L = load()
    n = len(views)
    arr = (type(views[0]) * n)(*views)
    tf = np.ascontiguousarray(tf, np.float64).reshape(-1)
    grid = np.ascontiguousarray(grid, np.float64).reshape(-1)
    gd = np.array(gdims, np.uint32)
    img = np.zeros(height * width * 4, np.float64)
    dtf = dgrid = None
    args = [None, None, None]
    if dimg is not None:
        dimg = np.ascontiguousarray(dimg, np.float64).reshape(-1)
        dtf = np.zeros(1024, np.float64)
        dgrid = np.zeros(grid.size, np.float64)
        args = [dimg.ctypes.data, dtf.ctypes.data, dgrid.ctypes.data]
    if window is None:
        rc = L.r3do_raycast_diff(C.byref(ovol.c), tf, grid, gd, C.cast(arr, VP), n, alpha_cutoff, img, width, height,
                                 *args)
    else:
        rc = L.r3do_raycast_diff_window(C.byref(ovol.c), tf, grid, gd, C.cast(arr, VP), n, alpha_cutoff, img, width,
                                        height, *args, *[int(x) for x in window])
    if rc != 0:
        raise ValueError("oracle raycast_diff: unsupported view")
    return img.reshape(height, width, 4), (None if dtf is None else dtf.reshape(256, 4)), dgrid


REF_DCT3D = os.path.join(HERE, "_ref", "libdct3d_ref.so")


def reference_decoder():
    """dct3d_decode_u8 of the reference codec built in oracle/_ref (the
    reference's own code, -O2 -ffp-contract=off), or None where it was not
    built. Test infrastructure: the checker and the timed CPU baseline only."""
    if not os.path.exists(REF_DCT3D):
        return None
    lib = C.CDLL(REF_DCT3D)
    fn = lib.dct3d_decode_u8
    fn.restype = C.c_int
    fn.argtypes = [VP, C.c_size_t, VP]
    reference_decoder.lib = lib  # keep it loaded
    return fn


def time_decoder(fn, data, offsets, lengths, iters, nthreads):
    """r3do_time_decoder over blobs (uint8 data, offsets, lengths) ->
    (seconds, successful decodes)."""
    data = np.ascontiguousarray(data, np.uint8)
    offs = np.ascontiguousarray(offsets, np.uint64)
    lens = np.ascontiguousarray(lengths, np.uint32)
    ok, sink = C.c_uint64(), C.c_uint64()
    dt = load().r3do_time_decoder(C.cast(fn, VP), data, offs.ctypes.data, lens.ctypes.data, lens.size, iters, nthreads,
                                  C.byref(ok), C.byref(sink))
    return dt, ok.value
