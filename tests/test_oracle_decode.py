"""The decode oracle pinned against the reference codec (golden vectors made
from tools/3ddct/dct3d.c, see tests/golden/make_golden.py) and against what
the mathematics fixes (orthonormal basis, DC-only and single-coefficient
closed forms, FP64 brute force, round trips)."""
import os

import numpy as np
import pytest

from oracle import oracle as orc

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "dct3d_golden.npz"))
L1 = np.add.outer(np.add.outer(np.arange(16), np.arange(16)), np.arange(16)).reshape(-1)


def fnv1a64(b):
    h = 0xCBF29CE484222325
    for byte in np.ascontiguousarray(b).tobytes():
        h = ((h ^ byte) * 0x100000001B3) & 0xFFFFFFFFFFFFFFFF
    return h


def test_basis_bit_exact_with_reference():
    assert np.array_equal(orc.basis().view(np.uint32), GOLD["basis"].view(np.uint32))


def test_basis_orthonormal_and_closed_form():
    cm = orc.basis().astype(np.float64).reshape(16, 16)
    assert np.allclose(cm @ cm.T, np.eye(16), atol=1e-6)
    n = np.arange(16)
    for k in range(16):
        ck = np.sqrt((1 if k == 0 else 2) / 16)
        assert np.allclose(cm[k], ck * np.cos(np.pi * (2 * n + 1) * k / 32), atol=1e-7)


def test_step_table_qstep_matches_reference():
    for i in range(len(GOLD["blob_names"])):
        q = np.float32(GOLD["blob_hdr"][i, 2])
        assert np.array_equal(orc.step_qstep(q).view(np.uint32), GOLD["blob_steps"][i].view(np.uint32))


@pytest.mark.parametrize("i", range(40))
def test_blob_decode_bit_exact(i):
    name = GOLD["blob_names"][i]
    c0, c1 = GOLD["blob_corr_off"][i], GOLD["blob_corr_off"][i + 1]
    ok, u8, pre = orc.decode_chunk(GOLD["blob_levels"][i], GOLD["blob_hdr"][i, 0], GOLD["blob_hdr"][i, 1],
                                   GOLD["blob_dcq"][i], GOLD["blob_steps"][i], GOLD["blob_corr_pos"][c0:c1],
                                   GOLD["blob_corr_delta"][c0:c1])
    assert ok == 1, name
    assert np.array_equal(pre.view(np.uint32), GOLD["blob_pre"][i].view(np.uint32)), name
    assert np.array_equal(u8, GOLD["blob_u8"][i]), name


def test_golden_set_covers_ties_corrections_and_q_range():
    assert float(np.min(GOLD["blob_min_tie"])) == 0.0  # an exact .5 is in the set
    assert int(GOLD["blob_corr_off"][-1]) > 0            # tau / max_error corrections
    qs = set(np.round(GOLD["blob_hdr"][:, 2], 3).tolist())
    assert {1.0, 2.0, 4.0, 8.0, 16.0, 32.0, 64.0} <= qs


def test_linear_table_vectors_bit_exact():
    step = GOLD["lin_step"]
    assert np.array_equal(step, orc.step_linear())
    for i in range(GOLD["lin_levels"].shape[0]):
        ok, u8, pre = orc.decode_chunk(GOLD["lin_levels"][i], GOLD["lin_vmin"][i], GOLD["lin_vspan"][i],
                                       GOLD["lin_dcq"][i], step)
        assert ok == 1
        assert np.array_equal(pre.view(np.uint32), GOLD["lin_pre"][i].view(np.uint32))
        assert np.array_equal(u8, GOLD["lin_u8"][i])


def test_hashes_of_512_sheet_chunks_match_reference():
    vol = orc.synth_sheets(128, 128, 128, 6, 1234)
    assert fnv1a64(vol) == int(GOLD["hash_volume_fnv"][0])
    step = orc.step_linear()
    lv, hdr = orc.encode_volume(vol, 128, 128, 128, step)
    assert fnv1a64(lv) == int(GOLD["hash_levels_fnv"][0])
    u8, status = orc.decode_batch(lv, hdr, step)
    assert status.all()
    got = np.array([fnv1a64(u8[i]) for i in range(u8.shape[0])], np.uint64)
    assert np.array_equal(got, GOLD["hash_fnv"])


def test_invalid_headers_rejected():
    lv = np.zeros(4096, np.int16)
    step = orc.step_linear()
    assert orc.decode_chunk(lv, 0.0, 0.0, 0, step)[0] == 0        # vspan 0
    assert orc.decode_chunk(lv, np.nan, 1.0, 0, step)[0] == 0     # vmin NaN
    assert orc.decode_chunk(lv, np.inf, 1.0, 0, step)[0] == 0
    assert orc.decode_chunk(lv, 0.0, -3.0, 0, step)[0] == 0


def test_dc_only_chunk_is_constant():
    for dcq in (0, 5, 1600, 4080, -16):
        lv = np.zeros(4096, np.int16)
        ok, u8, pre = orc.decode_chunk(lv, 0.0, 255.0, dcq, orc.step_linear())
        assert ok
        want = np.float32(np.float32(dcq) / np.float32(16)) * np.float32(1.0) + np.float32(0)
        assert np.all(pre == want)
        assert np.all(u8 == np.clip(np.floor(abs(want) + 0.5) * np.sign(want), 0, 255))


def test_dc_coefficient_closed_form():
    lv = np.zeros(4096, np.int16)
    lv[0] = 7
    step = orc.step_linear()
    ok, _, pre = orc.decode_chunk(lv, 0.0, 255.0, 0, step)
    coef = (7 - 1) * 1.0 + 0.8 + 0.4
    assert np.allclose(pre, coef / 64.0, atol=1e-5)


@pytest.mark.parametrize("pos", [(0, 0, 3), (2, 5, 0), (7, 1, 11), (15, 15, 15)])
def test_single_coefficient_is_separable_cosine_product(pos):
    kz, ky, kx = pos
    lv = np.zeros(4096, np.int16)
    lv[(kz * 16 + ky) * 16 + kx] = -9
    step = orc.step_linear()
    st = step[kz + ky + kx]
    coef = -((9 - 1) * float(st) + 0.8 * float(st) + 0.4 * float(st))
    ok, _, pre = orc.decode_chunk(lv, 0.0, 255.0, 32, step)
    n = np.arange(16)

    def c(k):
        ck = np.sqrt((1 if k == 0 else 2) / 16)
        return ck * np.cos(np.pi * (2 * n + 1) * k / 32)

    want = coef * np.einsum("z,y,x->zyx", c(kz), c(ky), c(kx)).reshape(-1) + 2.0
    assert np.max(np.abs(pre - want)) < 1e-3 * max(1.0, abs(coef) / 16)


def test_idct3_against_fp64_brute_force():
    rng = np.random.default_rng(3)
    coef = (rng.normal(0, 40, 4096) * (rng.random(4096) < 0.2)).astype(np.float32)
    out = np.zeros(4096, np.float32)
    orc.load().r3do_idct3(coef, out)
    cm = orc.basis().astype(np.float64).reshape(16, 16)
    want = np.einsum("az,by,cx,abc->zyx", cm, cm, cm, coef.astype(np.float64).reshape(16, 16, 16))
    assert np.max(np.abs(out.astype(np.float64) - want.reshape(-1))) < 1e-3


def test_forward_inverse_round_trip():
    rng = np.random.default_rng(5)
    blk = rng.uniform(-100, 100, 4096).astype(np.float32)
    coef = np.zeros(4096, np.float32)
    back = np.zeros(4096, np.float32)
    orc.load().r3do_fdct3(blk, coef)
    orc.load().r3do_idct3(coef, back)
    assert np.max(np.abs(back - blk)) < 1e-3


def test_a_planted_mistake_fails_the_golden_check():
    # drop the 0.4 term of deq_one: the reference vectors must reject it
    i = 10
    lv = GOLD["blob_levels"][i].astype(np.float32)
    st = GOLD["blob_steps"][i][L1]
    a = np.abs(lv)
    r = np.where(lv == 0, 0, np.sign(lv) * ((a - 1) * st + np.float32(0.8) * st)).astype(np.float32)
    out = np.zeros(4096, np.float32)
    orc.load().r3do_idct3(r, out)
    assert not np.array_equal((out + np.float32(GOLD["blob_dcq"][i] / 16)).view(np.uint32),
                              GOLD["blob_pre"][i].view(np.uint32))


def test_round_to_u8_as_one_biased_add():
    """The decoder rounds to u8 as trunc(clamp(v + c, 0, 255)) with
    c = 0.49999997f, the float just below 0.5, and the sum rounded to nearest
    even in FP32. That equals the reference's clamp(roundf(v), 0, 255) with
    NaN -> 0 (dct3d.c:625-635): here on every float within 4,096 ulps of each
    tie k + 0.5 (k = 0 .. 300) and of 0, on 200,000 random floats in
    [-2, 300], and on NaN, the infinities and the largest floats."""
    c = np.float32(0.49999997)
    assert c == np.nextafter(np.float32(0.5), np.float32(0)) and c < np.float32(0.5)

    def reference(v):
        w = v.astype(np.float64)
        r = np.where(np.abs(w) >= 2.0 ** 23, w, np.trunc(w + np.copysign(0.5, w)))
        return np.clip(np.where(np.isnan(w), 0.0, r), 0, 255).astype(np.uint32)

    def biased(v):
        s = (v + c).astype(np.float32)
        s = np.minimum(np.where(np.isnan(s), np.float32(0), np.maximum(s, np.float32(0))), np.float32(255))
        return np.trunc(s).astype(np.uint32)

    centres = np.concatenate([np.arange(301, dtype=np.float32) + np.float32(0.5), np.float32([0.0, 255.0])])
    bits = centres.view(np.uint32).astype(np.int64)[:, None] + np.arange(-4096, 4097)[None, :]
    near = bits[bits >= 0].astype(np.uint32).view(np.float32)
    rng = np.random.default_rng(3)
    rand = rng.uniform(-2.0, 300.0, 200_000).astype(np.float32)
    special = np.float32([np.nan, np.inf, -np.inf, 3.4e38, -3.4e38, -0.0, 1e-45, 0.49999997, 0.5, 255.5, 256.0])
    v = np.concatenate([near, -near[near < 2], rand, special])
    with np.errstate(invalid="ignore", over="ignore"):
        assert np.array_equal(reference(v), biased(v))
