"""Write shard-container vectors from the reference writer and reader.

Run where ``make -C oracle ref`` has built ``oracle/_ref/libshard_ref.so``
(src/shard/shardio.c, tools/3ddct/tools/export/shard.c and tools/3ddct/dct3d.c
compiled where they lie). Builds a store of dims (1024, 1536, 2048): shards
0_0_0, 0_0_1 and 0_1_0 are written by the reference ``shard_encode_u8``
(quality 8, tau 16: lodpack's level-0 settings, tools/lodpack/main.c:392-401)
from volumes that are zero except for groups of smooth 16^3 chunks placed
around the shard boundaries and the volume's y edge; shard 0_1_1 is absent.
Then the reference ``r3d_shard_decode_region`` decodes a list of regions
(aligned, straddling, strided, across shards, into the absent shard, past the
volume's edge, outside it), and damaged copies of shard 0_0_0 (bad crc, an
overflowing index entry, an invalid half-missing entry, corrupt payload
bytes, truncation). Writes ``dct3d_shards.npz`` (data only): the shard bytes,
the regions with the reference's return codes and bytes, and the damage
patches with theirs. Scratch files go under ``_work/`` of the repository.
"""
import ctypes as C
import os
import shutil
import sys

import numpy as np

from make_golden import smooth_block

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = C.CDLL(os.path.join(ROOT, "oracle", "_ref", "libshard_ref.so"))
LIBC = C.CDLL(None)
REF.shard_encode_u8.restype = C.c_int
REF.shard_encode_u8.argtypes = [C.c_void_p, C.c_float, C.c_float, C.POINTER(C.c_void_p), C.POINTER(C.c_size_t)]
REF.shard_crc32c.restype = C.c_uint32
REF.shard_crc32c.argtypes = [C.c_void_p, C.c_size_t]


class Store(C.Structure):
    _fields_ = [("dir", C.c_char * 512), ("nz", C.c_uint64), ("ny", C.c_uint64), ("nx", C.c_uint64)]


REF.r3d_shard_store_init.argtypes = [C.POINTER(Store), C.c_char_p, C.c_uint64, C.c_uint64, C.c_uint64]
REF.r3d_shard_decode_region.restype = C.c_int
REF.r3d_shard_decode_region.argtypes = [C.POINTER(Store)] + [C.c_uint64] * 3 + [C.c_uint32] * 3 + [C.c_void_p,
                                                                                                  C.c_int]
DIMS = (1024, 1536, 2048)
INDEX_BYTES = 64 ** 3 * 16 + 4
SCRATCH = os.path.join(ROOT, "_work", "shard_golden")


def encode_shard(chunks, rng):
    """chunks: shard-local (cz, cy, cx) with content; the rest of the 1024^3
    volume is zero (missing chunks)."""
    vol = np.zeros((1024, 1024, 1024), np.uint8)
    for cz, cy, cx in chunks:
        vol[cz * 16:cz * 16 + 16, cy * 16:cy * 16 + 16, cx * 16:cx * 16 + 16] = smooth_block(rng).reshape(16, 16, 16)
    out, n = C.c_void_p(), C.c_size_t()
    assert REF.shard_encode_u8(vol.ctypes.data, 8.0, 16.0, C.byref(out), C.byref(n)) == 0
    b = np.ctypeslib.as_array(C.cast(out, C.POINTER(C.c_uint8)), shape=(n.value,)).copy()
    LIBC.free(out)
    return b


def write_store(path, shards):
    shutil.rmtree(path, ignore_errors=True)
    os.makedirs(path)
    for (sz, sy, sx), b in shards.items():
        b.tofile(os.path.join(path, f"{sz}_{sy}_{sx}.shard"))
    st = Store()
    REF.r3d_shard_store_init(C.byref(st), path.encode(), *DIMS)
    return st


def decode(st, region):
    z0, y0, x0, dz, dy, dx = region
    dst = np.full(dz * dy * dx, 0xA5, np.uint8)
    rc = REF.r3d_shard_decode_region(C.byref(st), z0, y0, x0, dz, dy, dx, dst.ctypes.data, 4)
    return rc, dst


def main():
    rng = np.random.default_rng(11)
    groups = {
        (0, 0, 0): [(z, y, x) for z in range(30, 34) for y in range(16, 20) for x in range(60, 64)]
        + [(0, 0, 0), (63, 63, 63), (31, 63, 63)],
        (0, 0, 1): [(z, y, x) for z in range(30, 34) for y in range(16, 20) for x in range(0, 4)]
        + [(31, 63, 0), (0, 0, 63)],
        (0, 1, 0): [(z, y, x) for z in range(30, 34) for y in range(0, 4) for x in range(60, 64)]
        + [(31, y, x) for y in (31, 32) for x in (61, 62, 63)],  # y 1520..1551: past the volume's edge at 1536
    }
    for key in groups:  # plus scattered chunks
        groups[key] += [tuple(int(v) for v in rng.integers(0, 64, 3)) for _ in range(24)]
    shards = {k: encode_shard(sorted(set(v)), rng) for k, v in groups.items()}
    st = write_store(os.path.join(SCRATCH, "good"), shards)
    regions = [
        (480, 256, 960, 16, 16, 16),      # one chunk, aligned
        (488, 264, 968, 16, 16, 16),      # straddles 8 chunks
        (480, 256, 960, 8, 8, 32),        # wider than a chunk in x: row stride 32
        (483, 261, 1000, 50, 60, 60),     # across the x shard boundary
        (480, 1000, 990, 40, 600, 80),    # across y and x shards, into the absent shard, past y = 1536
        (496, 1500, 976, 16, 100, 48),    # past the volume's y edge only
        (1024, 0, 0, 8, 8, 8),            # outside the volume
        (487, 259, 971, 1, 1, 1),         # one voxel
        (0, 0, 768, 64, 320, 512),        # large, across shards
        (496, 1008, 1008, 48, 48, 48),    # the corner of four shards
        (0, 0, 0, 1024, 16, 16),          # a column of 64 chunks
    ]
    reg_rc, reg_off, reg_bytes = [], [0], []
    for r in regions:
        rc, dst = decode(st, r)
        reg_rc.append(rc)
        reg_bytes.append(dst)
        reg_off.append(reg_off[-1] + dst.size)
    # damaged copies of shard 0_0_0
    s0 = shards[(0, 0, 0)]
    idx0 = s0.size - INDEX_BYTES

    def entry(cz, cy, cx):
        return idx0 + ((cz * 64 + cy) * 64 + cx) * 16

    def with_crc(b):
        b = b.copy()
        crc = REF.shard_crc32c(b[idx0:].ctypes.data, INDEX_BYTES - 4)
        b[-4:] = np.array([crc], "<u4").view(np.uint8)
        return b

    def blob_of(cz, cy, cx):
        e = s0[entry(cz, cy, cx):entry(cz, cy, cx) + 16].view("<u8")
        return int(e[0]), int(e[1])

    damages = []
    b = s0.copy()
    b[-1] ^= 1
    damages.append(("bad crc", b, (480, 256, 960, 16, 16, 16)))
    b = s0.copy()
    b[entry(30, 16, 60):entry(30, 16, 60) + 16] = np.array([2 ** 64 - 8, 16], "<u8").view(np.uint8)
    damages.append(("overflowing entry, valid crc", with_crc(b), (488, 264, 968, 16, 16, 16)))
    b = s0.copy()
    b[entry(31, 17, 61):entry(31, 17, 61) + 16] = np.array([2 ** 64 - 1, 5], "<u8").view(np.uint8)
    damages.append(("half-missing entry, valid crc", with_crc(b), (488, 264, 968, 32, 32, 32)))
    b = s0.copy()
    off, n = blob_of(30, 17, 61)
    b[off + 2:off + min(n, 12)] ^= 0x5A
    damages.append(("corrupt payload", b, (480, 272, 976, 16, 16, 16)))
    b = s0.copy()
    off, n = blob_of(31, 16, 62)
    b[off] = 0x00  # magic byte
    damages.append(("bad magic", b, (496, 256, 992, 16, 16, 16)))
    damages.append(("truncated", s0[:INDEX_BYTES - 1].copy(), (480, 256, 960, 16, 16, 16)))
    dmg_patch_off, dmg_patch_len, dmg_patch_bytes, dmg_len = [], [], [], []
    dmg_rc, dmg_off, dmg_bytes, dmg_region = [], [0], [], []
    for k, (name, b, r) in enumerate(damages):
        st_d = write_store(os.path.join(SCRATCH, f"damage{k}"), {**shards, (0, 0, 0): b})
        rc, dst = decode(st_d, r)
        # the damage as a patch against s0: first and last differing byte, and the length
        m = min(b.size, s0.size)
        diff = np.nonzero(b[:m] != s0[:m])[0]
        p0, p1 = (int(diff[0]), int(diff[-1]) + 1) if diff.size else (0, 0)
        dmg_patch_off.append(p0)
        dmg_patch_len.append(p1 - p0)
        dmg_patch_bytes.append(b[p0:p1])
        dmg_len.append(b.size)
        dmg_rc.append(rc)
        dmg_bytes.append(dst)
        dmg_off.append(dmg_off[-1] + dst.size)
        dmg_region.append(r)
        print(f"{name}: rc {rc}")
    crc_data = rng.integers(0, 256, 1 << 16).astype(np.uint8)
    crc_lens = np.array([0, 1, 3, 4, 7, 64, 1000, 4097, 1 << 16], np.int64)
    crc_vals = np.array([REF.shard_crc32c(crc_data.ctypes.data, int(n)) for n in crc_lens], np.uint32)
    keys = sorted(shards)
    out = os.path.join(HERE, "dct3d_shards.npz")
    np.savez_compressed(
        out, dims=np.array(DIMS, np.uint64), shard_keys=np.array(keys, np.uint32),
        shard_offsets=np.cumsum([0] + [shards[k].size for k in keys]).astype(np.int64),
        shard_bytes=np.concatenate([shards[k] for k in keys]),
        regions=np.array(regions, np.uint64), region_rc=np.array(reg_rc, np.int32),
        region_offsets=np.array(reg_off, np.int64), region_bytes=np.concatenate(reg_bytes),
        damage_names=np.array([d[0] for d in damages]), damage_patch_off=np.array(dmg_patch_off, np.int64),
        damage_patch_len=np.array(dmg_patch_len, np.int64), damage_patch_bytes=np.concatenate(dmg_patch_bytes),
        damage_len=np.array(dmg_len, np.int64), damage_regions=np.array(dmg_region, np.uint64),
        damage_rc=np.array(dmg_rc, np.int32), damage_offsets=np.array(dmg_off, np.int64),
        damage_bytes=np.concatenate(dmg_bytes), crc_data=crc_data, crc_lens=crc_lens, crc_vals=crc_vals)
    shutil.rmtree(SCRATCH, ignore_errors=True)
    print("wrote", out, os.path.getsize(out), "bytes; region rcs", reg_rc)


if __name__ == "__main__":
    sys.exit(main())
