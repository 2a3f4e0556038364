"""Write the 3ddct golden vectors from the reference codec build.

Run in the container that holds the reference tree, after ``make -C oracle ref``
has compiled ``tools/3ddct/dct3d.c`` (with -ffp-contract=off) into
``oracle/_ref/libdct3d_ref.so``. The output, ``dct3d_golden.npz``, holds data
only: inputs and the reference's outputs at each stage of the decode back half.

Three sets:

* ``blob_*``: chunks encoded by the reference encoder (``dct3d_encode_u8``,
  dct3d.c:611) at q in {1..64}, some with tau / max_error corrections, some
  flat / DC-only / single-AC / extreme / near-tie. For each: the parsed header,
  the 46-entry step table, the levels, the correction list, the reference's
  pre-round FP32 values (``decode_float``, dct3d.c:542-593) and the u8 output
  of ``dct3d_decode_u8`` (dct3d.c:639).
* ``lin_*``: chunks quantised with the table q(i,j,k) = 1 + 2(i+j+k); the
  expected values come from the reference's own ``deq_one`` and ``dct3_inv``
  (dct3d.c:149-154, 113-118) followed by the decode_float epilogue.
* ``hash_*``: FNV-1a-64 of the reference-decoded u8 output of 4096 chunks of
  the seeded sheet volume, encoded with the oracle's encoder (the test
  regenerates the same levels and checks the hash).
"""
import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = C.CDLL(os.path.join(ROOT, "oracle", "_ref", "libdct3d_ref.so"))
ORA = C.CDLL(os.path.join(ROOT, "oracle", "liboracle.so"))

f32p = np.ctypeslib.ndpointer(np.float32, flags="C")
u8p = np.ctypeslib.ndpointer(np.uint8, flags="C")
i16p = np.ctypeslib.ndpointer(np.int16, flags="C")
u16p = np.ctypeslib.ndpointer(np.uint16, flags="C")

REF.dct3d_encode_u8.restype = C.c_size_t
REF.dct3d_encode_u8.argtypes = [u8p, C.c_float, C.c_float, C.c_float, u8p]
REF.dct3d_decode_u8.restype = C.c_int
REF.dct3d_decode_u8.argtypes = [u8p, C.c_size_t, u8p]
REF.ref_parse_u8.restype = C.c_int
REF.ref_parse_u8.argtypes = [u8p, C.c_size_t, C.POINTER(C.c_float), C.POINTER(C.c_float),
                             C.POINTER(C.c_float), C.POINTER(C.c_int32), f32p,
                             C.POINTER(C.c_uint32), u16p, f32p]
REF.ref_decode_float_u8.restype = C.c_int
REF.ref_decode_float_u8.argtypes = [u8p, C.c_size_t, f32p]
REF.ref_step_build.argtypes = [C.c_float, f32p]
REF.ref_basis.argtypes = [f32p]
REF.ref_deq_one.restype = C.c_float
REF.ref_deq_one.argtypes = [C.c_float, C.c_float]
REF.ref_idct3.argtypes = [f32p, f32p]

MAX_BYTES = 4096 * 8 + 256


def l1_index():
    z, y, x = np.meshgrid(np.arange(16), np.arange(16), np.arange(16), indexing="ij")
    return (z + y + x).reshape(-1)


L1 = l1_index()


def step46_from_full(step4096):
    t = np.zeros(46, np.float32)
    for f in range(46):
        t[f] = step4096[np.argmax(L1 == f)]
    return t


def roundf(v):
    """C roundf (half away from zero) on float32, exactly."""
    v = v.astype(np.float32)
    r = np.trunc(v)
    frac = (v - r).astype(np.float32)
    r = np.where(np.abs(frac) >= np.float32(0.5), r + np.sign(v), r)
    return r.astype(np.float32)


def finish_u8(rb, dcq, vmin, vspan, corr_pos=(), corr_delta=()):
    """decode_float epilogue (dct3d.c:568-591) + DEC_BODY (625-635) in float32."""
    dc = np.float32(np.float32(dcq) / np.float32(16.0))
    norm = (rb.astype(np.float32) + dc).astype(np.float32)
    for p, d in zip(corr_pos, corr_delta):
        norm[p] = np.float32(norm[p] + np.float32(d))
    inv = np.float32(np.float32(vspan) / np.float32(255.0))
    pre = (norm * inv).astype(np.float32) + np.float32(vmin)
    pre = pre.astype(np.float32)
    r = roundf(pre)
    r = np.where(np.isnan(r) | (r < 0), np.float32(0), r)
    r = np.minimum(r, np.float32(255))
    return pre, r.astype(np.uint8)


def smooth_block(rng, amp=60.0, noise=6.0, base=None):
    z, y, x = np.meshgrid(np.arange(16), np.arange(16), np.arange(16), indexing="ij")
    b = rng.uniform(60, 190) if base is None else base
    k = rng.uniform(0.05, 0.5, size=3)
    ph = rng.uniform(0, 6.28, size=3)
    v = b + amp * np.sin(k[0] * x + ph[0]) * np.cos(k[1] * y + ph[1]) + 0.5 * amp * np.sin(k[2] * z + ph[2])
    v = v + rng.normal(0, noise, v.shape)
    return np.clip(np.rint(v), 0, 255).astype(np.uint8).reshape(-1)


def special_blocks(rng):
    out = []
    out.append(("flat", np.full(4096, 137, np.uint8)))
    out.append(("flat0", np.zeros(4096, np.uint8)))
    out.append(("flat255", np.full(4096, 255, np.uint8)))
    zz, yy, xx = np.meshgrid(np.arange(16), np.arange(16), np.arange(16), indexing="ij")
    chk = (((zz + yy + xx) & 1) * 255).astype(np.uint8).reshape(-1)
    out.append(("checker", chk))
    ramp = (xx * 17).astype(np.uint8).reshape(-1)
    out.append(("ramp_x", ramp))
    cosz = np.clip(np.rint(128 + 100 * np.cos(np.pi * (2 * zz + 1) * 3 / 32)), 0, 255).astype(np.uint8).reshape(-1)
    out.append(("single_ac", cosz))
    out.append(("noise", rng.integers(0, 256, 4096).astype(np.uint8)))
    out.append(("sparse_spikes", np.where(rng.random(4096) < 0.01, 250, 10).astype(np.uint8)))
    return out


def encode(chunk, q, max_error=0.0, tau=0.0):
    buf = np.zeros(MAX_BYTES, np.uint8)
    n = REF.dct3d_encode_u8(np.ascontiguousarray(chunk), q, max_error, tau, buf)
    return buf[:n].copy()


def parse(blob):
    vmin, vspan, qstep = C.c_float(), C.c_float(), C.c_float()
    dcq = C.c_int32()
    lvl = np.zeros(4096, np.float32)
    nc = C.c_uint32()
    cpos = np.zeros(4096, np.uint16)
    cdel = np.zeros(4096, np.float32)
    ok = REF.ref_parse_u8(blob, blob.size, C.byref(vmin), C.byref(vspan), C.byref(qstep),
                          C.byref(dcq), lvl, C.byref(nc), cpos, cdel)
    assert ok
    assert np.all(np.abs(lvl) <= 32767)
    return (vmin.value, vspan.value, qstep.value, dcq.value, lvl.astype(np.int16),
            cpos[:nc.value].copy(), cdel[:nc.value].copy())


def tie_distance(pre):
    fr = np.abs(pre - np.trunc(pre))
    return float(np.min(np.abs(fr - 0.5)))


def make_blob_set(rng):
    cases = []
    for name, ch in special_blocks(rng):
        for q in (1.0, 8.0):
            cases.append((f"{name}_q{int(q)}", ch, q, 0.0, 0.0))
    for q in (1.0, 2.0, 4.0, 8.0, 16.0, 32.0, 64.0):
        for j in range(2):
            cases.append((f"smooth_q{int(q)}_{j}", smooth_block(rng), q, 0.0, 0.0))
    cases.append(("tau2_q4", smooth_block(rng), 4.0, 0.0, 2.0))
    cases.append(("tau1_q16", smooth_block(rng, noise=12), 16.0, 0.0, 1.0))
    cases.append(("maxerr_q8", smooth_block(rng), 8.0, 0.02, 0.0))
    cases.append(("qfrac_q2.5", smooth_block(rng), 2.5, 0.0, 0.0))
    # near-tie search: keep the blocks whose pre-round values sit closest to .5
    best = []
    for j in range(600):
        q = float(2 ** (j % 7))
        ch = smooth_block(rng)
        blob = encode(ch, q)
        pre = np.zeros(4096, np.float32)
        assert REF.ref_decode_float_u8(blob, blob.size, pre)
        best.append((tie_distance(pre), j, q, ch))
    best.sort(key=lambda t: t[0])
    for d, j, q, ch in best[:6]:
        cases.append((f"neartie_{j}_q{int(q)}", ch, q, 0.0, 0.0))

    names, blobs, offs = [], [], [0]
    hdr = np.zeros((len(cases), 4), np.float32)
    dcqs = np.zeros(len(cases), np.int32)
    steps = np.zeros((len(cases), 46), np.float32)
    lvls = np.zeros((len(cases), 4096), np.int16)
    pres = np.zeros((len(cases), 4096), np.float32)
    outs = np.zeros((len(cases), 4096), np.uint8)
    cpos_all, cdel_all, coffs = [], [], [0]
    min_tie = []
    for i, (name, ch, q, me, tau) in enumerate(cases):
        blob = encode(ch, q, me, tau)
        vmin, vspan, qstep, dcq, lvl, cpos, cdel = parse(blob)
        full = np.zeros(4096, np.float32)
        REF.ref_step_build(np.float32(qstep), full)
        pre = np.zeros(4096, np.float32)
        assert REF.ref_decode_float_u8(blob, blob.size, pre)
        out = np.zeros(4096, np.uint8)
        assert REF.dct3d_decode_u8(blob, blob.size, out) == 1
        names.append(name)
        blobs.append(blob)
        offs.append(offs[-1] + blob.size)
        hdr[i] = (vmin, vspan, qstep, 0)
        dcqs[i] = dcq
        steps[i] = step46_from_full(full)
        lvls[i] = lvl
        pres[i] = pre
        outs[i] = out
        cpos_all.append(cpos)
        cdel_all.append(cdel)
        coffs.append(coffs[-1] + cpos.size)
        min_tie.append(tie_distance(pre))
    return dict(
        blob_names=np.array(names), blob_bytes=np.concatenate(blobs), blob_offsets=np.array(offs, np.int64),
        blob_hdr=hdr, blob_dcq=dcqs, blob_steps=steps, blob_levels=lvls, blob_pre=pres, blob_u8=outs,
        blob_corr_pos=np.concatenate(cpos_all).astype(np.uint16),
        blob_corr_delta=np.concatenate(cdel_all).astype(np.float32),
        blob_corr_off=np.array(coffs, np.uint32), blob_min_tie=np.array(min_tie, np.float64),
    )


def ref_decode_levels(lvl, step46, dcq, vmin, vspan):
    coef = np.zeros(4096, np.float32)
    for i in np.nonzero(lvl)[0]:
        coef[i] = REF.ref_deq_one(float(lvl[i]), float(step46[L1[i]]))
    rb = np.zeros(4096, np.float32)
    REF.ref_idct3(coef, rb)
    return finish_u8(rb, dcq, vmin, vspan)


def oracle_encode_sheets(n=128, nsheets=6, seed=1234):
    vol = np.zeros(n * n * n, np.uint8)
    ORA.r3do_synth_sheets.argtypes = [u8p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                      C.c_uint64, C.c_uint32, C.c_uint32, C.c_int]
    ORA.r3do_synth_sheets(vol, n, n, n, nsheets, seed, 0, n, 8)
    step = np.zeros(46, np.float32)
    ORA.r3do_step_table_linear.argtypes = [f32p]
    ORA.r3do_step_table_linear(step)
    nch = (n // 16) ** 3
    lv = np.zeros(nch * 4096, np.int16)
    hdr = np.zeros(nch * 4, np.float32)  # r3do_chunk_hdr as 4x32-bit words
    ORA.r3do_encode_volume.argtypes = [u8p, C.c_uint32, C.c_uint32, C.c_uint32, f32p, i16p,
                                       C.c_void_p, C.c_int]
    ORA.r3do_encode_volume(vol, n, n, n, step, lv, hdr.ctypes.data, 8)
    dcq = hdr.view(np.int32).reshape(-1, 4)[:, 2].copy()
    h = hdr.reshape(-1, 4)
    return vol, step, lv.reshape(nch, 4096), h[:, 0].copy(), h[:, 1].copy(), dcq


def fnv1a64(b):
    h = 0xCBF29CE484222325
    for byte in b.tobytes():
        h ^= byte
        h = (h * 0x100000001B3) & 0xFFFFFFFFFFFFFFFF
    return h


def main():
    rng = np.random.default_rng(20261203)
    out = {}
    out.update(make_blob_set(rng))
    cm = np.zeros(256, np.float32)
    REF.ref_basis(cm)
    out["basis"] = cm

    vol, step, lv, vmin, vspan, dcq = oracle_encode_sheets()
    out["lin_step"] = step
    pick = rng.choice(lv.shape[0], 16, replace=False)
    lin_pre, lin_u8 = [], []
    for i in pick:
        pre, u8 = ref_decode_levels(lv[i], step, dcq[i], vmin[i], vspan[i])
        lin_pre.append(pre)
        lin_u8.append(u8)
    out.update(lin_levels=lv[pick], lin_vmin=vmin[pick], lin_vspan=vspan[pick], lin_dcq=dcq[pick],
               lin_pre=np.stack(lin_pre), lin_u8=np.stack(lin_u8), lin_index=pick.astype(np.int64))

    hashes = np.zeros(lv.shape[0], np.uint64)
    for i in range(lv.shape[0]):
        _, u8 = ref_decode_levels(lv[i], step, dcq[i], vmin[i], vspan[i])
        hashes[i] = fnv1a64(u8)
    out["hash_fnv"] = hashes
    out["hash_volume_fnv"] = np.array([fnv1a64(vol)], np.uint64)
    out["hash_levels_fnv"] = np.array([fnv1a64(lv)], np.uint64)
    path = os.path.join(HERE, "dct3d_golden.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes;", len(out["blob_names"]), "blobs;",
          "min tie distance", float(np.min(out["blob_min_tie"])))


if __name__ == "__main__":
    sys.exit(main())
