"""The oracle's 3ddct entropy stage against the reference's own blobs: the
reference parser's header, levels and corrections for all 40 golden blobs,
byte-identical re-encoding of the blobs that carry no corrections, and round
trips of generated levels."""
import os

import numpy as np

from oracle import oracle as orc
from render3d_amd import workloads

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "dct3d_golden.npz"))


def golden_blob(i):
    return GOLD["blob_bytes"][GOLD["blob_offsets"][i]:GOLD["blob_offsets"][i + 1]]


def test_parse_matches_the_reference_parser_on_every_golden_blob():
    for i in range(len(GOLD["blob_names"])):
        st, (vmin, vspan, qstep, dcq), lvl, cp, cd = orc.parse_blob(golden_blob(i))
        assert st == 1, GOLD["blob_names"][i]
        assert np.float32(vmin) == GOLD["blob_hdr"][i, 0] and np.float32(vspan) == GOLD["blob_hdr"][i, 1]
        assert np.float32(qstep) == GOLD["blob_hdr"][i, 2] and dcq == GOLD["blob_dcq"][i]
        assert np.array_equal(lvl, GOLD["blob_levels"][i]), GOLD["blob_names"][i]
        c0, c1 = GOLD["blob_corr_off"][i], GOLD["blob_corr_off"][i + 1]
        assert np.array_equal(cp, GOLD["blob_corr_pos"][c0:c1])
        assert np.array_equal(cd.view(np.uint32), GOLD["blob_corr_delta"][c0:c1].view(np.uint32))


def test_encode_reproduces_the_reference_blobs_bit_for_bit():
    n = 0
    for i in range(len(GOLD["blob_names"])):
        if GOLD["blob_corr_off"][i + 1] != GOLD["blob_corr_off"][i]:
            continue  # corrections are not written by the workload encoder
        h = GOLD["blob_hdr"][i]
        got = orc.encode_blob(h[0], h[1], h[2], int(GOLD["blob_dcq"][i]), GOLD["blob_levels"][i])
        assert np.array_equal(got, golden_blob(i)), GOLD["blob_names"][i]
        n += 1
    assert n >= 30


def test_round_trip_of_generated_levels():
    offs, pairs, hdr = workloads.sparse_laplacian_batch(64, 8, value_scale=300.0)
    lv = workloads.densify(offs, pairs)
    lv[0, :] = 0  # empty block
    lv[1, :] = 32767
    lv[2, ::5] = -32768
    rng = np.random.default_rng(1)
    for i in range(lv.shape[0]):
        vmin = float(np.float32(rng.uniform(-50, 50))) if i % 2 else float(rng.integers(-9, 9))
        vspan = float(np.float32(rng.uniform(1, 300)))
        qstep = float(np.float32(rng.choice([1.0, 2.5, 64.0, 0.3])))
        dcq = int(rng.integers(-5000, 5000))
        blob = orc.encode_blob(vmin, vspan, qstep, dcq, lv[i])
        st, h, got, cp, _ = orc.parse_blob(blob)
        assert st == 1 and cp.size == 0
        assert h == (np.float32(vmin), np.float32(vspan), np.float32(qstep), dcq)
        assert np.array_equal(got, lv[i]), i


def test_malformed_and_wrong_dtype_blobs_are_rejected():
    b = golden_blob(0).copy()
    assert orc.parse_blob(b, dtype=2)[0] == 0  # dtype mismatch
    b2 = b.copy()
    b2[0] = 0x00
    assert orc.parse_blob(b2)[0] == 0  # magic
    assert orc.parse_blob(b[:1])[0] == 0  # shorter than the preamble
    rng = np.random.default_rng(0)
    for _ in range(200):  # corrupt streams parse without reading past the blob
        c = b.copy()
        k = rng.integers(2, c.size)
        c[k] ^= np.uint8(1 << int(rng.integers(0, 8)))
        st = orc.parse_blob(c[: rng.integers(2, c.size + 1)])[0]
        assert st in (-1, 0, 1)


CORRUPT = np.load(os.path.join(os.path.dirname(__file__), "golden", "dct3d_corrupt.npz"))


def test_corrupt_blobs_parse_as_the_reference_parses_them():
    n = CORRUPT["ok"].size
    wide = 0
    for i in range(n):
        b = CORRUPT["bytes"][CORRUPT["offsets"][i]:CORRUPT["offsets"][i + 1]]
        st, (vmin, vspan, qstep, dcq), lvl, cp, cd = orc.parse_blob(b)
        if not CORRUPT["ok"][i]:
            assert st == 0, i
            continue
        ref = CORRUPT["levels"][i]
        if np.any(ref > 32767) or np.any(ref < -32768):
            assert st == -1, i  # levels the int16 formats cannot hold are reported, not clamped silently
            wide += 1
            continue
        assert st == 1, i
        assert np.array_equal(lvl.astype(np.float32), ref), i
        assert (np.float32(vmin), np.float32(vspan), np.float32(qstep)) == tuple(CORRUPT["hdr"][i]) and \
            dcq == CORRUPT["dcq"][i], i
        c0, c1 = CORRUPT["corr_off"][i], CORRUPT["corr_off"][i + 1]
        assert np.array_equal(cp, CORRUPT["corr_pos"][c0:c1]), i
        assert np.array_equal(cd.view(np.uint32), CORRUPT["corr_delta"][c0:c1].view(np.uint32)), i
    print(f"{int(CORRUPT['ok'].sum())} accepted by the reference, {wide} with levels beyond int16")


def test_blob_batch_decode_is_dct3d_decode_u8():
    offs = GOLD["blob_offsets"].astype(np.int64)
    u8, st = orc.decode_blobs(GOLD["blob_bytes"], offs[:-1], np.diff(offs), nthreads=4)
    assert np.all(st == 1)
    assert np.array_equal(u8, GOLD["blob_u8"])  # the reference decoder's output for the same blobs
    offs = CORRUPT["offsets"].astype(np.int64)
    u8, st = orc.decode_blobs(CORRUPT["bytes"], offs[:-1], np.diff(offs), nthreads=4)
    assert np.array_equal(st != 0, CORRUPT["ok"].astype(bool))
    assert not u8[st != 1].any()
