"""Write corrupted-blob vectors for the 3ddct entropy stage from the reference
parser build.

Run where ``make -C oracle ref`` has built ``oracle/_ref/libdct3d_ref.so``
(the reference ``tools/3ddct/dct3d.c`` plus the probe ``oracle/ref_probe.c``).
Reads the reference-encoded blobs of ``dct3d_golden.npz`` and writes
``dct3d_corrupt.npz``: 160 damaged variants (single bit flips, byte
overwrites, truncations), plus 8 blobs the reference encoder writes at tiny
quality (0.003-0.1) whose levels exceed int16, and, for each, what the
reference parser
(``ref_parse_u8``: the entropy part of decode_float, dct3d.c:542-587)
returns: the accept flag, the header, the levels (as the reference's floats)
and the corrections. Data only.
"""
import os
import sys

import numpy as np

from make_golden import HERE, REF, encode, u16p  # noqa: F401  (REF carries the argtypes)
import ctypes as C


def parse(blob):
    vmin, vspan, qstep = C.c_float(), C.c_float(), C.c_float()
    dcq = C.c_int32()
    lvl = np.zeros(4096, np.float32)
    nc = C.c_uint32()
    cpos = np.zeros(4096, np.uint16)
    cdel = np.zeros(4096, np.float32)
    ok = REF.ref_parse_u8(blob if blob.size else np.zeros(1, np.uint8), blob.size, C.byref(vmin), C.byref(vspan),
                          C.byref(qstep), C.byref(dcq), lvl, C.byref(nc), cpos, cdel)
    return ok, (vmin.value, vspan.value, qstep.value, dcq.value), lvl, cpos[:nc.value].copy(), cdel[:nc.value].copy()


def main():
    g = np.load(os.path.join(HERE, "dct3d_golden.npz"))
    rng = np.random.default_rng(7)
    blobs, offs = [], [0]
    oks, hdrs, dcqs, lvls, coffs, cps, cds, src = [], [], [], [], [0], [], [], []
    n = len(g["blob_names"])
    for k in range(160):
        i = int(rng.integers(0, n))
        b = g["blob_bytes"][g["blob_offsets"][i]:g["blob_offsets"][i + 1]].copy()
        kind = k % 4
        if kind == 0:  # one bit flip in the range-coded stream
            j = int(rng.integers(2, b.size))
            b[j] ^= np.uint8(1 << int(rng.integers(0, 8)))
        elif kind == 1:  # a run of overwritten bytes
            j = int(rng.integers(2, b.size))
            m = int(rng.integers(1, 9))
            b[j:j + m] = rng.integers(0, 256, min(m, b.size - j)).astype(np.uint8)
        elif kind == 2:  # truncated
            b = b[: int(rng.integers(2, b.size))]
        else:  # flip early, in the header fields
            j = int(rng.integers(2, min(b.size, 10)))
            b[j] ^= np.uint8(1 << int(rng.integers(0, 8)))
        ok, (vmin, vspan, qstep, dcq), lvl, cp, cd = parse(b)
        blobs.append(b)
        offs.append(offs[-1] + b.size)
        oks.append(ok)
        hdrs.append((vmin, vspan, qstep))
        dcqs.append(dcq)
        lvls.append(lvl)
        cps.append(cp)
        cds.append(cd)
        coffs.append(coffs[-1] + cp.size)
        src.append(i)
    # valid blobs whose levels exceed int16: the reference keeps levels as
    # floats; the int16 level formats must report them rather than wrap
    zz, yy, xx = np.meshgrid(np.arange(16), np.arange(16), np.arange(16), indexing="ij")
    chk = (((zz + yy + xx) & 1) * 255).astype(np.uint8).reshape(-1)
    ramp = (xx * 17).astype(np.uint8).reshape(-1)
    for q in (0.003, 0.01, 0.03, 0.1):
        for blk in (chk, ramp):
            b = encode(blk, q)
            ok, (vmin, vspan, qstep, dcq), lvl, cp, cd = parse(b)
            assert ok == 1
            blobs.append(b)
            offs.append(offs[-1] + b.size)
            oks.append(ok)
            hdrs.append((vmin, vspan, qstep))
            dcqs.append(dcq)
            lvls.append(lvl)
            cps.append(cp)
            cds.append(cd)
            coffs.append(coffs[-1] + cp.size)
            src.append(-1)
    path = os.path.join(HERE, "dct3d_corrupt.npz")
    np.savez_compressed(path, bytes=np.concatenate(blobs), offsets=np.array(offs, np.int64),
                        ok=np.array(oks, np.int32), hdr=np.array(hdrs, np.float32), dcq=np.array(dcqs, np.int32),
                        levels=np.stack(lvls), corr_off=np.array(coffs, np.uint32),
                        corr_pos=np.concatenate(cps).astype(np.uint16),
                        corr_delta=np.concatenate(cds).astype(np.float32), source=np.array(src, np.int32))
    wide = int(sum(np.any(np.abs(v) > 32767) for v in lvls))
    print("wrote", path, os.path.getsize(path), "bytes;", int(np.sum(oks)), "of", len(oks), "accepted;", wide,
          "with levels beyond int16")


if __name__ == "__main__":
    sys.exit(main())
