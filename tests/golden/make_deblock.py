"""Write deblocking vectors from the reference deblock.c.

Run where ``make -C oracle ref`` has built ``oracle/_ref/libdeblock_ref.so``
(tools/3ddct/deblock.c) and ``oracle/_ref/libdct3d_ref.so`` (the codec, for
the blocky inputs). Inputs are decoded volumes with seams: the smooth ripple
of tools/3ddct/test_deblock.c:82-90 sent through the reference codec at
q = 24 (64^3), a second ripple at q = 32 cropped to 40 x 50 x 70, the
test's genuine-edge volume and a 17 x 3 x 33 sliver. Outputs: the reference
dct3d_deblock_u8 for several (step, strength) pairs, and dct3d_deblock_chunk_u8
for chunks of the 64^3 volume with every neighbour pattern. Data only:
writes ``dct3d_deblock.npz``.
"""
import ctypes as C
import os
import sys

import numpy as np

from make_golden import REF, encode

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
DEB = C.CDLL(os.path.join(ROOT, "oracle", "_ref", "libdeblock_ref.so"))
DEB.dct3d_deblock_u8.argtypes = [C.c_void_p, C.c_void_p, C.c_float, C.c_float]
DEB.dct3d_deblock_chunk_u8.argtypes = [C.c_void_p] * 7 + [C.c_float, C.c_float, C.c_void_p]


def roundtrip(vol, q):
    """Encode and decode every 16^3 block with the reference codec."""
    nz, ny, nx = vol.shape
    out = np.zeros_like(vol)
    for z in range(0, nz, 16):
        for y in range(0, ny, 16):
            for x in range(0, nx, 16):
                blob = encode(np.ascontiguousarray(vol[z:z + 16, y:y + 16, x:x + 16]).reshape(-1), q)
                back = np.zeros(4096, np.uint8)
                assert REF.dct3d_decode_u8(blob, blob.size, back) == 1
                out[z:z + 16, y:y + 16, x:x + 16] = back.reshape(16, 16, 16)
    return out


def ripple(nz, ny, nx, a=0.11, b=0.09, c=0.07):
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    f = 128.0 + 60.0 * np.sin(x * a + 0.5) * np.cos(y * b) + 40.0 * np.sin(z * c + y * 0.03)
    return (np.clip(f, 0, 255) + 0.5).astype(np.uint8)  # test_deblock.c:84-89


def deblock(vol, step, strength):
    v = np.ascontiguousarray(vol).copy()
    dims = (C.c_size_t * 3)(*v.shape)
    DEB.dct3d_deblock_u8(v.ctypes.data, dims, step, strength)
    return v


def main():
    a = roundtrip(ripple(64, 64, 64), 24.0)
    b = roundtrip(ripple(48, 64, 80, 0.13, 0.05, 0.09), 32.0)[:40, :50, :70].copy()
    edge = np.where(np.arange(64)[None, None, :] < 32, 40, 200).astype(np.uint8) * np.ones((64, 64, 1), np.uint8)
    sliver = roundtrip(ripple(32, 16, 48), 40.0)[5:22, 7:10, 2:35].copy()
    vols = {"a": a, "b": b, "edge": edge, "sliver": sliver}
    cases = [("a", 0.0, 1.0), ("a", 24.0, 1.0), ("a", 24.0, 0.0), ("b", 32.0, 0.5), ("b", 32.0, 2.0),
             ("b", 0.0, 1.0), ("edge", 0.0, 1.0), ("edge", 24.0, 1.0), ("sliver", 40.0, 1.0), ("sliver", 0.0, 1.5)]
    outs = [deblock(vols[v], s, t) for v, s, t in cases]
    # chunk API over the 4 x 4 x 4 chunks of a
    chunks = a.reshape(4, 16, 4, 16, 4, 16).transpose(0, 2, 4, 1, 3, 5).reshape(64, 4096).copy()
    rng = np.random.default_rng(3)
    ck_center, ck_mask, ck_step, ck_str, ck_out = [], [], [], [], []
    for k in range(24):
        c = int(rng.integers(0, 64))
        cz, cy, cx = c // 16, (c // 4) % 4, c % 4
        nb = [(cz, cy, cx - 1), (cz, cy, cx + 1), (cz, cy - 1, cx), (cz, cy + 1, cx), (cz - 1, cy, cx), (cz + 1, cy, cx)]
        mask = int(rng.integers(0, 64)) if k % 3 else 63
        ptrs = []
        for f, (z, y, x) in enumerate(nb):
            ok = (mask >> f) & 1 and 0 <= z < 4 and 0 <= y < 4 and 0 <= x < 4
            ptrs.append(chunks[(z * 4 + y) * 4 + x].ctypes.data if ok else None)
            if not ok:
                mask &= ~(1 << f)
        step = (0.0, 24.0, 12.0)[k % 3]
        strength = (1.0, 0.7, 1.6, 1.0)[k % 4]
        out = np.zeros(4096, np.uint8)
        DEB.dct3d_deblock_chunk_u8(chunks[c].ctypes.data, *ptrs, step, strength, out.ctypes.data)
        ck_center.append(c)
        ck_mask.append(mask)
        ck_step.append(step)
        ck_str.append(strength)
        ck_out.append(out)
    path = os.path.join(HERE, "dct3d_deblock.npz")
    np.savez_compressed(path, a=a, b=b, sliver=sliver, case_vol=np.array([c[0] for c in cases]),
                        case_step=np.array([c[1] for c in cases], np.float32),
                        case_strength=np.array([c[2] for c in cases], np.float32),
                        case_out=np.concatenate([o.reshape(-1) for o in outs]),
                        case_off=np.cumsum([0] + [o.size for o in outs]).astype(np.int64),
                        ck_center=np.array(ck_center, np.int32), ck_mask=np.array(ck_mask, np.int32),
                        ck_step=np.array(ck_step, np.float32), ck_strength=np.array(ck_str, np.float32),
                        ck_out=np.stack(ck_out))
    changed = [int((o != vols[v]).sum()) for o, (v, _, _) in zip(outs, cases)]
    print("wrote", path, os.path.getsize(path), "bytes; voxels changed per case", changed)


if __name__ == "__main__":
    sys.exit(main())
