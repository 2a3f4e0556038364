"""Zarr v3 shards on the GPU through the C ABI: the index crc32c against the
reference writer's values and the published check value; region decodes of
shards the reference wrote, compared byte for byte with the reference
reader's r3d_shard_decode_region (tests/golden/dct3d_shards.npz), including
damaged shards and the safety cases of tests/test_shard_safety.c; per-chunk
statuses against the oracle; and, at full size, a GPU-built 1024^3 shard whose
region decode equals the direct decode of its levels."""
import os

import numpy as np
import pytest

from oracle import oracle as orc
from render3d_amd import R3DError, workloads
from render3d_amd import _lib

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

G = np.load(os.path.join(os.path.dirname(__file__), "golden", "dct3d_shards.npz"))
DIMS = tuple(int(v) for v in G["dims"])


def cuda(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host_shards():
    out = []
    for i, (sz, sy, sx) in enumerate(G["shard_keys"]):
        out.append((G["shard_bytes"][G["shard_offsets"][i]:G["shard_offsets"][i + 1]], int(sz), int(sy), int(sx)))
    return out


def dev_shards(hs=None):
    return [(cuda(b), sz, sy, sx) for b, sz, sy, sx in (hs or host_shards())]


def damaged(k):
    s0 = host_shards()[0][0].copy()
    p0, pl = int(G["damage_patch_off"][k]), int(G["damage_patch_len"][k])
    s0[p0:p0 + pl] = G["damage_patch_bytes"][int(G["damage_patch_len"][:k].sum()):int(G["damage_patch_len"][:k + 1].sum())]
    return s0[:int(G["damage_len"][k])]


def test_crc32c(ctx):
    assert ctx.crc32c(cuda(np.frombuffer(b"123456789", np.uint8))) == 0xE3069283
    for n, v in zip(G["crc_lens"], G["crc_vals"]):
        if n:
            assert ctx.crc32c(cuda(G["crc_data"][:n])) == v, n
    rng = np.random.default_rng(5)
    big = rng.integers(0, 256, (1 << 22) + 7).astype(np.uint8)
    d = cuda(big)
    for a, b in ((0, big.size), (1, big.size - 2), (3, 5000), (2, 1025), (0, 1024), (1, 1)):
        assert ctx.crc32c(d[a:a + b]) == orc.crc32c(big[a:a + b]), (a, b)


def test_regions_match_the_reference_reader(ctx):
    ds = dev_shards()
    for i, r in enumerate(G["regions"]):
        r = tuple(int(v) for v in r)
        print("region", r, flush=True)
        dst, counts, st = ctx.shard_decode_region(ds, DIMS, r)
        torch.cuda.synchronize()
        want = G["region_bytes"][G["region_offsets"][i]:G["region_offsets"][i + 1]]
        assert np.array_equal(dst.cpu().numpy().reshape(-1), want), r
        c = counts.cpu().numpy()
        assert c[2] == 0 and c[3] == 0 and G["region_rc"][i] == 0
        orc_rc, _, orc_counts, orc_st = orc.shard_decode_region(host_shards(), DIMS, r)
        assert np.array_equal(c, orc_counts.astype(np.int32)), r
        n = ctx.shard_region_chunks(DIMS, r)
        assert n == orc_st.size or (n == 0 and r[0] >= DIMS[0])
        if n:
            assert np.array_equal(st.cpu().numpy()[:n], orc_st), r


def test_damaged_shards_match_the_reference_reader(ctx):
    rest = dev_shards()[1:]
    for k, name in enumerate(G["damage_names"]):
        r = tuple(int(v) for v in G["damage_regions"][k])
        print("damage", name, r, flush=True)
        dst, counts, st = ctx.shard_decode_region([(cuda(damaged(k)), 0, 0, 0)] + rest, DIMS, r)
        torch.cuda.synchronize()
        want = G["damage_bytes"][G["damage_offsets"][k]:G["damage_offsets"][k + 1]]
        assert np.array_equal(dst.cpu().numpy().reshape(-1), want), name
        assert counts.cpu().numpy()[2] > 0 and G["damage_rc"][k] == -1, name  # the reference returned -1
        _, _, _, orc_st = orc.shard_decode_region([(damaged(k), 0, 0, 0)] + host_shards()[1:], DIMS, r)
        assert np.array_equal(st.cpu().numpy()[:orc_st.size], orc_st), name


def test_safety_cases(ctx):
    """tests/test_shard_safety.c:57-90, and argument checks."""
    dims = (1024, 1024, 1024)
    empty = orc.shard_pack({})
    dst, counts, st = ctx.shard_decode_region([(cuda(empty), 0, 0, 0)], dims, (0, 0, 0, 16, 16, 16))
    torch.cuda.synchronize()
    assert not dst.any() and counts.cpu().numpy().tolist() == [0, 1, 0, 0] and st.cpu().numpy()[0] == 0
    bad = empty.copy()
    bad[-4] ^= 1
    dst, counts, st = ctx.shard_decode_region([(cuda(bad), 0, 0, 0)], dims, (0, 0, 0, 16, 16, 16))
    torch.cuda.synchronize()
    assert counts.cpu().numpy()[2] == 1 and st.cpu().numpy()[0] == -1
    over = empty.copy()
    e = over.size - orc.SHARD_INDEX_BYTES
    over[e:e + 16] = np.array([2 ** 64 - 8, 16], "<u8").view(np.uint8)
    over[-4:] = np.array([orc.crc32c(over[e:-4])], "<u4").view(np.uint8)
    dst, counts, st = ctx.shard_decode_region([(cuda(over), 0, 0, 0)], dims, (0, 0, 0, 16, 16, 16))
    torch.cuda.synchronize()
    assert counts.cpu().numpy()[2] == 1 and not dst.any()
    short = cuda(empty[:100])  # shorter than its index: corrupt
    dst, counts, _ = ctx.shard_decode_region([(short, 0, 0, 0)], dims, (0, 0, 0, 16, 16, 32))
    torch.cuda.synchronize()
    assert counts.cpu().numpy()[2] == 2
    big = 2 ** 32 - 1
    with pytest.raises(R3DError):
        ctx.shard_decode_region([], dims, (0, 0, 0, big, big, big), out=torch.empty(1, dtype=torch.uint8,
                                                                                     device="cuda"), status=False)
    with pytest.raises(R3DError):
        ctx.shard_decode_region([], dims, (0, 0, 0, 0, 4, 4), out=torch.empty(1, dtype=torch.uint8, device="cuda"),
                                status=False)
    # outside the volume: zeros, no chunk
    dst, counts, _ = ctx.shard_decode_region([], dims, (2000, 0, 0, 4, 4, 4), status=False)
    torch.cuda.synchronize()
    assert not dst.any() and not counts.any()


def test_shard_store_reads_files(ctx, tmp_path):
    for b, sz, sy, sx in host_shards():
        b.tofile(tmp_path / f"{sz}_{sy}_{sx}.shard")
    from render3d_amd import ShardStore
    store = ShardStore(ctx, str(tmp_path), *DIMS)
    for i in (1, 4, 9):
        r = tuple(int(v) for v in G["regions"][i])
        rc, dst = store.decode_region(*r)
        want = G["region_bytes"][G["region_offsets"][i]:G["region_offsets"][i + 1]]
        assert rc == 0 and np.array_equal(dst.cpu().numpy().reshape(-1), want), r
    assert store.shard(0, 1, 1) is None  # absent


def test_full_shard_equals_level_decode(ctx):
    """A 1024^3 shard built on the GPU (sheet volume, qstep-8 levels, blobs,
    index, crc) with every 7th chunk left missing: the whole-shard region
    decode equals the decode of the same levels, and missing chunks are 0."""
    n = 1024
    step = torch.from_numpy(workloads.step_table_qstep(8.0)).cuda()
    vol = ctx.synth_sheets(n, n, n, 48, 99)
    levels, headers = ctx.encode_chunks(vol, n, n, n, step)
    del vol
    data, offs, lens = ctx.encode_blobs(levels, headers, 8.0)
    present = torch.arange(64 ** 3, device="cuda") % 7 != 3
    shard = workloads.pack_shard(ctx, data, offs, lens, present)
    del data
    flat = ctx.decode_chunks(headers=headers, step_tables=step, levels=levels)
    del levels
    flat[~present] = 0
    want = flat.view(64, 64, 64, 16, 16, 16).permute(0, 3, 1, 4, 2, 5).reshape(n, n, n)
    dst, counts, st = ctx.shard_decode_region([(shard, 0, 0, 0)], (n, n, n), (0, 0, 0, n, n, n))
    torch.cuda.synchronize()
    assert torch.equal(dst, want)
    c = counts.cpu().numpy()
    assert c[1] == int((~present).sum()) and c[0] == int(present.sum()) and c[2] == 0 and c[3] == 0
    assert torch.equal(st == 1, present)
    # an unaligned sub-region of the same shard
    r = (100, 333, 517, 200, 150, 301)
    dst, counts, _ = ctx.shard_decode_region([(shard, 0, 0, 0)], (n, n, n), r, status=False)
    torch.cuda.synchronize()
    assert torch.equal(dst, want[100:300, 333:483, 517:818])


def test_volume_decode_shards_equals_region_then_upload(ctx):
    """r3d_hip_volume_decode_shards: the region decode straight into a render
    volume equals the region decode uploaded into one (every level and the
    occupancy), with the same counts, on the reference's shards (aligned,
    straddling shards, past the volume's y edge, odd sizes)."""
    ds = dev_shards()
    for i in (0, 3, 4, 5, 8, 9):
        z0, y0, x0, dz, dy, dx = (int(v) for v in G["regions"][i])
        v = ctx.volume(dx, dy, dz)
        counts = v.decode_shards(ds, DIMS, (z0, y0, x0))
        dst, c2, _ = ctx.shard_decode_region(ds, DIMS, (z0, y0, x0, dz, dy, dx), status=False)
        w = ctx.volume(dx, dy, dz)
        w.upload(dst.contiguous())
        torch.cuda.synchronize()
        assert np.array_equal(counts.cpu().numpy(), c2.cpu().numpy()), i
        for l in range(v.nlevels):
            assert np.array_equal(v.read_level(l), w.read_level(l)), (i, l)
        assert np.array_equal(v.read_occupancy(), w.read_occupancy()), i
        assert np.array_equal(v.read_level(0), dst.cpu().numpy().reshape(-1)), i
    with pytest.raises(R3DError):
        v = ctx.volume(16, 16, 16)
        r = ctx._region(DIMS, (0, 0, 0, 16, 16, 32))  # the region must be the volume's size
        import ctypes
        ctx.check(ctx.lib.r3d_hip_volume_decode_shards(ctx.h, v.h, None, 0, ctypes.byref(r), None, None, None))
