"""The oracle's shard container (oracle/oracle_shard.c) against the reference
reader and writer: shards written by the reference shard_encode_u8
(tools/3ddct/tools/export/shard.c), region decodes and damaged shards decoded
by the reference r3d_shard_decode_region (src/shard/shardio.c), all in
tests/golden/dct3d_shards.npz; plus the safety cases of
tests/test_shard_safety.c and the published CRC-32C check value."""
import os

import numpy as np

from oracle import oracle as orc

G = np.load(os.path.join(os.path.dirname(__file__), "golden", "dct3d_shards.npz"))
DIMS = tuple(int(v) for v in G["dims"])


def shards():
    out = []
    for i, (sz, sy, sx) in enumerate(G["shard_keys"]):
        out.append((G["shard_bytes"][G["shard_offsets"][i]:G["shard_offsets"][i + 1]], int(sz), int(sy), int(sx)))
    return out


def damaged(k):
    s0 = shards()[0][0]
    assert tuple(G["shard_keys"][0]) == (0, 0, 0)
    b = s0.copy()
    p0, pl = int(G["damage_patch_off"][k]), int(G["damage_patch_len"][k])
    pb = G["damage_patch_bytes"][int(G["damage_patch_len"][:k].sum()):int(G["damage_patch_len"][:k + 1].sum())]
    b[p0:p0 + pl] = pb
    return b[:int(G["damage_len"][k])]


def test_crc32c_check_value_and_reference():
    assert orc.crc32c(np.frombuffer(b"123456789", np.uint8)) == 0xE3069283  # the published CRC-32C check value
    for n, v in zip(G["crc_lens"], G["crc_vals"]):
        assert orc.crc32c(G["crc_data"][:n]) == v, n  # the reference writer's shard_crc32c


def test_reference_shards_are_valid_and_repack_byte_identical():
    for b, sz, sy, sx in shards():
        assert orc.shard_valid(b)
        idx = b[b.size - orc.SHARD_INDEX_BYTES:-4].view("<u8").reshape(-1, 2)
        present = np.nonzero(~((idx[:, 0] == 2 ** 64 - 1) & (idx[:, 1] == 2 ** 64 - 1)))[0]
        assert present.size > 50
        blobs = {int(ci): b[int(idx[ci, 0]):int(idx[ci, 0] + idx[ci, 1])] for ci in present}
        assert np.array_equal(orc.shard_pack(blobs), b)  # shard.c's layout: payload in chunk order, index, crc


def test_regions_match_the_reference():
    sh = shards()
    for i, r in enumerate(G["regions"]):
        r = tuple(int(v) for v in r)
        rc, dst, counts, _ = orc.shard_decode_region(sh, DIMS, r)
        want = G["region_bytes"][G["region_offsets"][i]:G["region_offsets"][i + 1]]
        assert rc == G["region_rc"][i], r
        assert np.array_equal(dst.reshape(-1), want), r
        assert counts[2] == 0 and counts[3] == 0
    # the region into the absent shard and past the volume has missing chunks and zeros
    rc, dst, counts, st = orc.shard_decode_region(sh, DIMS, (480, 1000, 990, 40, 600, 80))
    assert counts[1] > 0 and np.all(st >= 0) and not dst[:, 1536 - 1000:, :].any()


def test_damaged_shards_match_the_reference():
    sh = shards()
    for k, name in enumerate(G["damage_names"]):
        r = tuple(int(v) for v in G["damage_regions"][k])
        rc, dst, counts, _ = orc.shard_decode_region([(damaged(k), 0, 0, 0)] + sh[1:], DIMS, r)
        want = G["damage_bytes"][G["damage_offsets"][k]:G["damage_offsets"][k + 1]]
        assert rc == G["damage_rc"][k] == -1, name
        assert counts[2] > 0, name
        assert np.array_equal(dst.reshape(-1), want), name


def test_safety_cases_of_the_reference_tests():
    """tests/test_shard_safety.c:57-90."""
    empty = orc.shard_pack({})  # every entry missing
    rc, dst, counts, _ = orc.shard_decode_region([(empty, 0, 0, 0)], (1024, 1024, 1024), (0, 0, 0, 16, 16, 16))
    assert rc == 0 and not dst.any() and counts[1] == 1
    bad = empty.copy()
    bad[-4] ^= 1
    assert not orc.shard_valid(bad)
    rc, _, counts, _ = orc.shard_decode_region([(bad, 0, 0, 0)], (1024, 1024, 1024), (0, 0, 0, 16, 16, 16))
    assert rc == -1 and counts[2] == 1
    over = empty.copy()
    e = over.size - orc.SHARD_INDEX_BYTES
    over[e:e + 16] = np.array([2 ** 64 - 8, 16], "<u8").view(np.uint8)
    over[-4:] = np.array([orc.crc32c(over[e:-4])], "<u4").view(np.uint8)
    assert orc.shard_valid(over)  # a valid crc does not legitimise an overflowing range
    rc, dst, _, st = orc.shard_decode_region([(over, 0, 0, 0)], (1024, 1024, 1024), (0, 0, 0, 16, 16, 16))
    assert rc == -1 and st[0] == -1 and not dst.any()
    big = 2 ** 32 - 1
    import ctypes as C
    L = orc.load()
    arr = (orc.ShardC * 1)()
    dst = np.zeros(16, np.uint8)
    assert L.r3do_shard_decode_region(C.addressof(arr), 0, 1024, 1024, 1024, 0, 0, 0, big, big, big, dst, None, None,
                                      1) == -1


def test_planted_mistake_is_caught():
    """Decoding with the index read one entry off (a plausible indexing
    mistake) does not reproduce the reference's region."""
    b, *_ = shards()[0]
    idx0 = b.size - orc.SHARD_INDEX_BYTES
    shifted = b.copy()
    shifted[idx0:idx0 + orc.SHARD_GRID ** 3 * 16 - 16] = b[idx0 + 16:idx0 + orc.SHARD_GRID ** 3 * 16]
    shifted[-4:] = np.array([orc.crc32c(shifted[idx0:-4])], "<u4").view(np.uint8)
    r = tuple(int(v) for v in G["regions"][1])
    _, dst, _, _ = orc.shard_decode_region([(shifted, 0, 0, 0)] + shards()[1:], DIMS, r)
    assert not np.array_equal(dst.reshape(-1), G["region_bytes"][G["region_offsets"][1]:G["region_offsets"][2]])
