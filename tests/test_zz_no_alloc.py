"""r3d_hip_parse_blobs_noalloc through the C ABI: the allocation-free parse
(lane order and sort scratch in a fixed per-GPU device array) gives exactly
what r3d_hip_parse_blobs gives, on two calls in a row (the second smaller),
accepts the documented maximum batch and rejects one blob more before
launching anything."""
import ctypes as C
import os

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

HERE = os.path.dirname(__file__)
GOLD = np.load(os.path.join(HERE, "golden", "dct3d_golden.npz"))
CORRUPT = np.load(os.path.join(HERE, "golden", "dct3d_corrupt.npz"))
KEYS = ("headers", "step_tables", "pair_offsets", "pair_counts", "pairs", "corr_offsets", "corr_counts", "corr_pos",
        "corr_delta", "status")


def cuda(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def tiled(byts, offs, rep, zero_every=0):
    n = offs.size - 1
    lens1 = np.diff(offs).astype(np.int64)
    starts, lens = [], []
    for k in range(rep):
        for i in range(n):
            starts.append(offs[i] + k * byts.size)
            lens.append(lens1[i])
            if zero_every and (k * n + i) % zero_every == 0:
                starts.append(0)
                lens.append(0)
    return cuda(np.tile(byts, rep)), cuda(np.array(starts, np.int64)), cuda(np.array(lens, np.int32))


def outputs(n, cap, fill):
    """Output arrays pre-filled with a sentinel, so words a call does not
    write are seen as such."""
    o = {"headers": torch.full((n, 4), fill, dtype=torch.int32, device="cuda"),
         "step_tables": torch.full((n * 46,), fill, dtype=torch.int32, device="cuda"),
         "pair_offsets": torch.full((n + 1,), fill, dtype=torch.int32, device="cuda"),
         "pair_counts": torch.full((n,), fill, dtype=torch.int32, device="cuda"),
         "pairs": torch.full((n * cap,), fill, dtype=torch.int32, device="cuda"),
         "corr_offsets": torch.full((n + 1,), fill, dtype=torch.int32, device="cuda"),
         "corr_counts": torch.full((n,), fill, dtype=torch.int32, device="cuda"),
         "corr_pos": torch.full((n * cap,), fill, dtype=torch.int16, device="cuda"),
         "corr_delta": torch.full((n * cap,), fill, dtype=torch.int32, device="cuda"),
         "status": torch.full((n,), fill, dtype=torch.int32, device="cuda")}
    return o


def call(ctx, name, data, starts, lens, cap, fill):
    from render3d_amd._lib import BlobBatch, BlobParseOut, R3D_DT_U8
    n = int(starts.shape[0])
    o = outputs(n, cap, fill)
    bb = BlobBatch(n, R3D_DT_U8, cap, 0, data.data_ptr(), starts.data_ptr(), lens.data_ptr())
    po = BlobParseOut(*[o[k].data_ptr() for k in KEYS])
    rc = getattr(ctx.lib, name)(ctx.h, C.byref(bb), C.byref(po), None)
    torch.cuda.synchronize()
    return rc, {k: v.cpu().numpy() for k, v in o.items()}


def same(a, b):
    """Equal wherever the parse writes: counts bound the pair and correction
    slots a chunk fills, the rest of its capacity is not written."""
    for k in ("headers", "step_tables", "pair_offsets", "pair_counts", "corr_offsets", "corr_counts", "status"):
        assert np.array_equal(a[k], b[k]), k
    cap = a["pairs"].size // a["status"].size
    for c in range(a["status"].size):
        np_, nc = int(a["pair_counts"][c]), int(a["corr_counts"][c])
        sl = slice(c * cap, c * cap + np_)
        assert np.array_equal(a["pairs"][sl], b["pairs"][sl]), c
        sl = slice(c * cap, c * cap + nc)
        assert np.array_equal(a["corr_pos"][sl], b["corr_pos"][sl]) and \
            np.array_equal(a["corr_delta"][sl], b["corr_delta"][sl]), c


def test_noalloc_parse_equals_parse_twice_then_smaller(ctx):
    # first call: the reference's 40 blobs x 52 = 2,080 (length-ordered lanes)
    data, starts, lens = tiled(GOLD["blob_bytes"], GOLD["blob_offsets"].astype(np.int64), 52)
    rc, a = call(ctx, "r3d_hip_parse_blobs_noalloc", data, starts, lens, 4096, -7)
    assert rc == 0, ctx.lib.r3d_hip_last_error(ctx.h)
    rc, b = call(ctx, "r3d_hip_parse_blobs", data, starts, lens, 4096, -7)
    assert rc == 0
    same(a, b)
    assert np.all(a["status"] == 1)
    # second call, smaller: the damaged and low-quality reference blobs with
    # zero-length entries, 616 entries (still length-ordered)
    data, starts, lens = tiled(CORRUPT["bytes"], CORRUPT["offsets"].astype(np.int64), 3, zero_every=5)
    assert 512 <= starts.shape[0] < 2080
    rc, a = call(ctx, "r3d_hip_parse_blobs_noalloc", data, starts, lens, 4096, 5)
    assert rc == 0, ctx.lib.r3d_hip_last_error(ctx.h)
    rc, b = call(ctx, "r3d_hip_parse_blobs", data, starts, lens, 4096, 5)
    assert rc == 0
    same(a, b)
    assert {-1, 0, 1} <= set(a["status"].tolist())


def test_noalloc_parse_limits(ctx):
    from render3d_amd._lib import R3D_PARSE_NOALLOC_MAX_BLOBS
    # the largest batch: zero-length blobs (each rejected, status 0)
    n = R3D_PARSE_NOALLOC_MAX_BLOBS
    data = torch.zeros(16, dtype=torch.uint8, device="cuda")
    starts = torch.zeros(n + 1, dtype=torch.int64, device="cuda")
    lens = torch.zeros(n + 1, dtype=torch.int32, device="cuda")
    rc, a = call(ctx, "r3d_hip_parse_blobs_noalloc", data, starts[:n], lens[:n], 1, 9)
    assert rc == 0, ctx.lib.r3d_hip_last_error(ctx.h)
    assert np.all(a["status"] == 0) and np.all(a["pair_counts"] == 0)
    # one more: refused before any launch (the sentinel survives)
    rc, a = call(ctx, "r3d_hip_parse_blobs_noalloc", data, starts, lens, 1, 9)
    assert rc == -5  # R3D_ERR_RANGE
    assert b"262144" in ctx.lib.r3d_hip_last_error(ctx.h)
    assert np.all(a["status"] == 9)
