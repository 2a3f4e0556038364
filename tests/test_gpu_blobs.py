"""3ddct blobs on the GPU through the C ABI: the reference's own blobs parsed
and decoded end to end (bit-exact with dct3d_decode_u8), corrupted blobs
treated as the reference parser treats them, and the workload encoder
reproducing the reference encoder's bytes."""
import os

import numpy as np
import pytest

from oracle import oracle as orc
from render3d_amd import workloads

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

HERE = os.path.dirname(__file__)
GOLD = np.load(os.path.join(HERE, "golden", "dct3d_golden.npz"))
CORRUPT = np.load(os.path.join(HERE, "golden", "dct3d_corrupt.npz"))


def cuda(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def upload(byts, offsets):
    n = offsets.size - 1
    lengths = np.diff(offsets).astype(np.int32)
    return cuda(byts), cuda(offsets[:n].astype(np.int64)), cuda(lengths)


def levels_of(parsed, cap):
    pairs = parsed["pairs"].cpu().numpy().view(np.uint32)
    cnt = parsed["pair_counts"].cpu().numpy()
    out = np.zeros((cnt.size, 4096), np.int16)
    for c in range(cnt.size):
        p = pairs[c * cap:c * cap + cnt[c]]
        out[c, p & 4095] = (p >> 16).astype(np.uint16).view(np.int16)
    return out


def test_golden_blobs_parse_and_decode_bit_exact(ctx):
    n = len(GOLD["blob_names"])
    data, offs, lens = upload(GOLD["blob_bytes"], GOLD["blob_offsets"])
    parsed = ctx.parse_blobs(data, offs, lens)
    torch.cuda.synchronize()
    assert np.all(parsed["status"].cpu().numpy() == 1)
    hdr = parsed["headers"].cpu().numpy()
    assert np.array_equal(hdr[:, 0].view(np.float32), GOLD["blob_hdr"][:, 0])
    assert np.array_equal(hdr[:, 1].view(np.float32), GOLD["blob_hdr"][:, 1])
    assert np.array_equal(hdr[:, 2], GOLD["blob_dcq"]) and np.array_equal(hdr[:, 3], np.arange(n))
    steps = parsed["step_tables"].cpu().numpy().reshape(n, 46)
    assert np.array_equal(steps.view(np.uint32), GOLD["blob_steps"].view(np.uint32))  # qstep * powf(1+f, 0.65)
    assert np.array_equal(levels_of(parsed, 4096), GOLD["blob_levels"])
    cc = parsed["corr_counts"].cpu().numpy()
    cp = parsed["corr_pos"].cpu().numpy().view(np.uint16)
    cd = parsed["corr_delta"].cpu().numpy()
    for i in range(n):
        c0, c1 = GOLD["blob_corr_off"][i], GOLD["blob_corr_off"][i + 1]
        assert cc[i] == c1 - c0
        assert np.array_equal(cp[i * 4096:i * 4096 + cc[i]], GOLD["blob_corr_pos"][c0:c1])
        assert np.array_equal(cd[i * 4096:i * 4096 + cc[i]].view(np.uint32),
                              GOLD["blob_corr_delta"][c0:c1].view(np.uint32))
    u8, st = ctx.decode_blobs(data, offs, lens)
    torch.cuda.synchronize()
    bad = [GOLD["blob_names"][i] for i in range(n) if not np.array_equal(u8[i].cpu().numpy(), GOLD["blob_u8"][i])]
    assert not bad, f"differing from dct3d_decode_u8: {bad}"


def test_corrupt_blobs_follow_the_reference_parser(ctx):
    n = CORRUPT["ok"].size
    data, offs, lens = upload(CORRUPT["bytes"], CORRUPT["offsets"])
    parsed = ctx.parse_blobs(data, offs, lens)
    u8, _ = ctx.decode_blobs(data, offs, lens)
    torch.cuda.synchronize()
    st = parsed["status"].cpu().numpy()
    lv = levels_of(parsed, 4096)
    u8 = u8.cpu().numpy()
    wide = 0
    for i in range(n):
        if not CORRUPT["ok"][i]:
            assert st[i] == 0, i
            assert not u8[i].any(), i  # rejected: zero-filled
            continue
        ref = CORRUPT["levels"][i]
        if np.any(ref > 32767) or np.any(ref < -32768):
            assert st[i] == -1, i  # valid, but beyond the int16 level formats: reported, zero-filled
            assert not u8[i].any(), i
            wide += 1
            continue
        assert st[i] == 1, i
        assert np.array_equal(lv[i].astype(np.float32), CORRUPT["levels"][i]), i
        hdr = parsed["headers"][i].cpu().numpy()
        assert hdr[0:2].view(np.float32).tolist() == CORRUPT["hdr"][i, 0:2].tolist() and hdr[2] == CORRUPT["dcq"][i]
    assert wide >= 4


def test_encoder_reproduces_reference_blobs(ctx):
    idx = [i for i in range(len(GOLD["blob_names"])) if GOLD["blob_corr_off"][i + 1] == GOLD["blob_corr_off"][i]]
    hdr = np.zeros((len(idx), 4), np.int32)
    hdr[:, 0] = GOLD["blob_hdr"][idx, 0].view(np.int32)
    hdr[:, 1] = GOLD["blob_hdr"][idx, 1].view(np.int32)
    hdr[:, 2] = GOLD["blob_dcq"][idx]
    data, offs, lens = ctx.encode_blobs(cuda(GOLD["blob_levels"][idx]), cuda(hdr), cuda(GOLD["blob_hdr"][idx, 2]))
    torch.cuda.synchronize()
    data, offs, lens = data.cpu().numpy(), offs.cpu().numpy(), lens.cpu().numpy()
    for j, i in enumerate(idx):
        want = GOLD["blob_bytes"][GOLD["blob_offsets"][i]:GOLD["blob_offsets"][i + 1]]
        assert lens[j] == want.size, GOLD["blob_names"][i]
        assert np.array_equal(data[offs[j]:offs[j] + lens[j]], want), GOLD["blob_names"][i]


def test_volume_round_trip_and_oracle_encoder_agreement(ctx):
    n = 64
    vol = ctx.synth_sheets(n, n, n, 3, 77)
    qstep = np.float32(2.0)
    step = cuda(orc.step_qstep(qstep))  # qstep * powf(1 + f, 0.65f), the table the blob's qstep implies
    levels, headers = ctx.encode_chunks(vol, n, n, n, step)
    data, offs, lens = ctx.encode_blobs(levels, headers, float(qstep))
    u8, st = ctx.decode_blobs(data, offs, lens)
    direct = ctx.decode_chunks(headers=headers, step_tables=step, levels=levels)
    torch.cuda.synchronize()
    assert np.all(st.cpu().numpy() == 1)
    assert torch.equal(u8, direct)
    lv, hd = levels.cpu().numpy(), headers.cpu().numpy()
    d, o, ln = data.cpu().numpy(), offs.cpu().numpy(), lens.cpu().numpy()
    for c in range(0, lv.shape[0], 5):
        want = orc.encode_blob(hd[c, 0:1].view(np.float32)[0], hd[c, 1:2].view(np.float32)[0], qstep, int(hd[c, 2]),
                               lv[c])
        assert np.array_equal(d[o[c]:o[c] + ln[c]], want), c


def test_pair_output_paths(ctx):
    """Pairs leave a lane eight at a time (32-byte groups) when the capacity is
    a multiple of 8, one at a time otherwise; both give the reference's levels
    for every golden blob, whatever its count modulo 8."""
    data, offs, lens = upload(GOLD["blob_bytes"], GOLD["blob_offsets"])
    counts = (GOLD["blob_levels"] != 0).sum(axis=1)
    assert len(set(int(c) % 8 for c in counts)) > 4
    cap8 = int(counts.max()) + 8 - int(counts.max()) % 8
    for cap in (cap8, cap8 + 1, cap8 + 4, cap8 + 8):
        parsed = ctx.parse_blobs(data, offs, lens, max_pairs=cap)
        torch.cuda.synchronize()
        assert np.all(parsed["status"].cpu().numpy() == 1), cap
        assert np.array_equal(parsed["pair_counts"].cpu().numpy(), counts), cap
        assert np.array_equal(levels_of(parsed, cap), GOLD["blob_levels"]), cap


def test_capacity_dtype_and_empty_blobs(ctx):
    offs_np, pairs, hdr = workloads.sparse_laplacian_batch(8, 3)
    lv = workloads.densify(offs_np, pairs)
    counts = np.diff(offs_np.astype(np.int64))
    data, offs, lens = ctx.encode_blobs(cuda(lv), cuda(hdr), 1.0)
    cap = int(np.sort(counts)[3])  # the larger half overflows
    parsed = ctx.parse_blobs(data, offs, lens, max_pairs=cap)
    u8, _ = ctx.decode_blobs(data, offs, lens, max_pairs=cap)
    torch.cuda.synchronize()
    st = parsed["status"].cpu().numpy()
    assert np.array_equal(st == -2, counts > cap) and np.all(st[counts <= cap] == 1)
    assert not u8[torch.from_numpy(counts > cap).cuda()].any()
    parsed = ctx.parse_blobs(data, offs, lens, dtype=2)  # blobs carry dtype 1
    torch.cuda.synchronize()
    assert np.all(parsed["status"].cpu().numpy() == 0)
    short = torch.zeros(8, dtype=torch.int32, device="cuda") + 1  # one byte: shorter than the preamble
    parsed = ctx.parse_blobs(data, offs, short)
    torch.cuda.synchronize()
    assert np.all(parsed["status"].cpu().numpy() == 0)


def test_parse_is_independent_of_blob_alignment(ctx):
    """The parser reads each blob through aligned dwords: the same blobs at
    every byte alignment, the last one ending at the buffer's end, parse alike."""
    byts, offs = GOLD["blob_bytes"], GOLD["blob_offsets"].astype(np.int64)
    n = offs.size - 1
    ref = None
    for shift in range(4):
        order = np.roll(np.arange(n), shift)  # a different blob ends the buffer each time
        parts = [np.zeros(shift, np.uint8)] + [byts[offs[i]:offs[i + 1]] for i in order]
        data = np.concatenate(parts)
        lens = np.diff(offs)[order]
        starts = shift + np.concatenate([[0], np.cumsum(lens)[:-1]])
        assert starts[-1] + lens[-1] == data.size
        parsed = ctx.parse_blobs(cuda(data), cuda(starts.astype(np.int64)), cuda(lens.astype(np.int32)))
        torch.cuda.synchronize()
        got = {k: parsed[k].cpu().numpy() for k in ("status", "headers", "pair_counts", "corr_counts")}
        lv = levels_of(parsed, 4096)
        inv = np.argsort(order)
        got = {k: v[inv] if k != "headers" else v[inv, :3] for k, v in got.items()}
        got["levels"] = lv[inv]
        assert np.all(got["status"] == 1)
        if ref is None:
            ref = got
            assert np.array_equal(got["levels"], GOLD["blob_levels"])
        for k in ref:
            assert np.array_equal(got[k], ref[k]), (shift, k)


def test_large_batches_parse_in_length_order(ctx):
    """Batches of 512 blobs or more are parsed longest blob first (the lanes
    take a radix-sorted order); the outputs stay indexed by blob. The
    reference's 40 blobs repeated 16 times, each copy at its own offset."""
    n = len(GOLD["blob_names"])
    byts, offs = GOLD["blob_bytes"], GOLD["blob_offsets"].astype(np.int64)
    rep = 16
    data = np.tile(byts, rep)
    big_offs = np.concatenate([offs[:n] + k * byts.size for k in range(rep)])
    lens = np.tile(np.diff(offs).astype(np.int32), rep)
    parsed = ctx.parse_blobs(cuda(data), cuda(big_offs), cuda(lens))
    torch.cuda.synchronize()
    assert np.all(parsed["status"].cpu().numpy() == 1)
    assert np.array_equal(levels_of(parsed, 4096), np.tile(GOLD["blob_levels"], (rep, 1)))
    hdr = parsed["headers"].cpu().numpy()
    assert np.array_equal(hdr[:, 2], np.tile(GOLD["blob_dcq"], rep)) and np.array_equal(hdr[:, 3], np.arange(n * rep))
    steps = parsed["step_tables"].cpu().numpy().reshape(n * rep, 46)
    assert np.array_equal(steps.view(np.uint32), np.tile(GOLD["blob_steps"].view(np.uint32), (rep, 1)))
    assert np.array_equal(parsed["corr_counts"].cpu().numpy(), np.tile(np.diff(GOLD["blob_corr_off"]), rep))
    po = parsed["pair_offsets"].cpu().numpy()
    assert np.array_equal(po, np.arange(n * rep + 1) * 4096)


def _tiled(byts, offs, rep, zero_every=0):
    """The blobs of (byts, offs) repeated rep times, each copy at its own
    offset; with zero_every > 0 a zero-length entry (a missing chunk, as the
    shard gather emits) follows every zero_every-th blob."""
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
    return np.tile(byts, rep), np.array(starts, np.int64), np.array(lens, np.int32)


def _parse_np(c, data, starts, lens, cap=4096):
    p = c.parse_blobs(cuda(data), cuda(starts), cuda(lens), max_pairs=cap)
    torch.cuda.synchronize()
    return {"status": p["status"].cpu().numpy(), "headers": p["headers"].cpu().numpy(),
            "pair_counts": p["pair_counts"].cpu().numpy(), "corr_counts": p["corr_counts"].cpu().numpy(),
            "levels": levels_of(p, cap)}


def test_damaged_blobs_in_length_order_match_the_unsorted_parse(ctx):
    """The 160 corrupted blobs and 8 low-quality blobs of the reference (with
    their verdicts) tiled past 512 entries with zero-length entries mixed in:
    the length-ordered parse gives every blob the status, header, counts and
    levels of the same blob parsed alone in a batch below 512 (chunk order)."""
    byts, offs = CORRUPT["bytes"], CORRUPT["offsets"].astype(np.int64)
    data, starts, lens = _tiled(byts, offs, 4, zero_every=7)
    assert starts.size >= 512 and np.sum(lens == 0) > 50
    big = _parse_np(ctx, data, starts, lens)
    ref = {k: [] for k in big}
    for a in range(0, starts.size, 300):  # batches below 512: lanes in chunk order
        part = _parse_np(ctx, data, starts[a:a + 300], lens[a:a + 300])
        for k in ref:
            ref[k].append(part[k])
    ref = {k: np.concatenate(v) for k, v in ref.items()}
    assert np.array_equal(big["headers"][:, 3], np.arange(starts.size))  # table = the chunk's index in its call
    big["headers"], ref["headers"] = big["headers"][:, :3], ref["headers"][:, :3]
    for k in ref:
        assert np.array_equal(big[k], ref[k]), k
    assert np.all(big["status"][lens == 0] == 0)
    assert {-1, 0, 1} <= set(big["status"].tolist())


def test_repeated_parses_agree_with_the_reference(ctx):
    """Every lane stages its pair groups in its own LDS halfwords only: the
    other waves of a workgroup are still writing their priors and EOB
    contexts in the same rows, at a timing that differs from run to run (a
    32-bit staging slot over two other lanes' halfwords once turned a -1
    level into +1). Repeated parses of mixed batches, lane-sorted and not,
    must all equal the reference's levels."""
    n = len(GOLD["blob_names"])
    byts, offs = GOLD["blob_bytes"], GOLD["blob_offsets"].astype(np.int64)
    for rep, zero_every in ((52, 0), (12, 3)):  # 2,080 entries in length order; 640 with zero-length entries mixed in
        data, starts, lens = _tiled(byts, offs, rep, zero_every)
        want = np.tile(GOLD["blob_levels"], (rep, 1))
        real = lens > 0
        for _ in range(8):
            got = _parse_np(ctx, data, starts, lens)
            assert np.all(got["status"][real] == 1) and np.all(got["status"][~real] == 0)
            assert np.array_equal(got["levels"][real], want)
    data, starts, lens = _tiled(byts, offs, 11)  # 440 entries: lanes in chunk order
    for _ in range(8):
        got = _parse_np(ctx, data, starts, lens)
        assert np.array_equal(got["levels"], np.tile(GOLD["blob_levels"], (11, 1)))


def test_captured_parse_survives_a_larger_eager_parse():
    """A HIP graph holding a length-ordered parse keeps its lane-order buffer
    when a later eager parse on the same context needs a larger one (the old
    buffer is retired until the context is destroyed), and a capture that
    would have to grow the buffer is refused before anything is recorded."""
    import render3d_amd as r3d

    c = r3d.Context(0)  # a fresh context: its order buffer starts empty
    try:
        byts, offs = GOLD["blob_bytes"], GOLD["blob_offsets"].astype(np.int64)
        ds, ss, ls = _tiled(byts, offs, 13)  # 520 blobs
        db, sb, lb = _tiled(byts, offs, 52)  # 2,080 blobs
        want = _parse_np(c, ds, ss, ls)
        d_data, d_st, d_ln = cuda(ds), cuda(ss), cuda(ls)
        g = torch.cuda.CUDAGraph()
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            with torch.cuda.graph(g, stream=s):
                gp = c.parse_blobs(d_data, d_st, d_ln, stream=s)
        torch.cuda.current_stream().wait_stream(s)
        torch.cuda.synchronize()
        big = _parse_np(c, db, sb, lb)  # the order buffer grows here
        assert np.all(big["status"] == 1)
        assert np.array_equal(big["levels"], np.tile(GOLD["blob_levels"], (52, 1)))
        for rep in range(3):
            for k in ("status", "pair_counts", "headers"):
                gp[k].zero_()
            gp["pairs"].zero_()
            junk = torch.randn(1 << 20, device="cuda").sum()  # eager work between replays
            _parse_np(c, db[:byts.size * 20], sb[:800], lb[:800])
            torch.cuda.synchronize()
            g.replay()
            torch.cuda.synchronize()
            assert np.array_equal(gp["status"].cpu().numpy(), want["status"]), rep
            assert np.array_equal(gp["pair_counts"].cpu().numpy(), want["pair_counts"]), rep
            assert np.array_equal(gp["headers"].cpu().numpy(), want["headers"]), rep
            assert np.array_equal(levels_of(gp, 4096), want["levels"]), rep
            del junk
        del g
        # a capture that would grow the buffer: refused, nothing launched
        dx, sx, lx = _tiled(byts, offs, 128)  # 5,120 blobs: more than the buffer holds
        d_dx, d_sx, d_lx = cuda(dx), cuda(sx), cuda(lx)
        g2 = torch.cuda.CUDAGraph()
        s.wait_stream(torch.cuda.current_stream())
        marker = torch.zeros(1, device="cuda")
        with pytest.raises(r3d.R3DError, match="capturing"):
            with torch.cuda.stream(s):
                with torch.cuda.graph(g2, stream=s):
                    marker.add_(1.0)  # the graph is not empty
                    c.parse_blobs(d_dx, d_sx, d_lx, stream=s)
        torch.cuda.synchronize()
        after = _parse_np(c, dx, sx, lx)  # eagerly the same batch grows the buffer and parses
        assert np.all(after["status"] == 1)
    finally:
        torch.cuda.synchronize()
        c.close()
