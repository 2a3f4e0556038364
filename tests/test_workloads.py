"""Synthetic batched workloads: scan order, sparse batch statistics, view packing."""
import numpy as np

from render3d_amd import workloads


def test_scan_order_is_ascending_l1_with_index_ties():
    sc = workloads.scan_order()
    assert sorted(sc.tolist()) == list(range(4096))
    l1 = sc // 256 + (sc // 16) % 16 + sc % 16
    assert np.all(np.diff(l1) >= 0)
    for f in (0, 7, 45):
        run = sc[l1 == f]
        assert np.all(np.diff(run.astype(int)) > 0)
    assert sc[0] == 0 and sc[-1] == 4095 and sc[1:4].tolist() == [1, 16, 256]


def test_sparse_batch_statistics_and_round_trip():
    offs, pairs, hdr = workloads.sparse_laplacian_batch(2000, 3)
    counts = np.diff(offs.astype(np.int64))
    assert counts.min() >= 5 and counts.max() <= 400 and 60 < counts.mean() < 140
    assert offs[-1] == pairs.size
    rank = np.empty(4096, np.int64)
    rank[workloads.scan_order()] = np.arange(4096)
    dense = workloads.densify(offs, pairs)
    assert np.array_equal((dense != 0).sum(1), counts)  # distinct positions, nonzero levels
    for c in (0, 1, 999):
        pos = pairs[offs[c]:offs[c + 1]] & 4095
        assert rank[pos].max() < 4 * counts[c]  # inside the low-L1 prefix
    lv = (pairs >> 16).astype(np.uint16).view(np.int16)
    assert np.all(lv != 0)
    assert 8 < np.abs(lv).mean() < 16  # Laplace(0, 12) magnitudes
    vmin, vspan = hdr[:, 0].view(np.float32), hdr[:, 1].view(np.float32)
    assert vmin.min() >= 0 and vspan.min() >= 120 and np.all(hdr[:, 3] == 0)
    again = workloads.sparse_laplacian_batch(2000, 3)
    assert all(np.array_equal(a, b) for a, b in zip((offs, pairs, hdr), again))


def test_mixed_views_tile_the_atlas_without_overlap():
    views, w, h = workloads.mixed_views()
    assert len(views) == 64 and w == 7680
    assert sum(v.viewport[0] * v.viewport[1] for v in views) == 53_166_080
    cover = np.zeros((h, w), np.uint8)
    for v in views:
        x, y = v.view_org & 0xFFFF, v.view_org >> 16
        assert x + v.viewport[0] <= w and y + v.viewport[1] <= h
        cover[y:y + v.viewport[1], x:x + v.viewport[0]] += 1
    assert cover.max() == 1


def test_qstep_table_matches_the_reference_blobs():
    import os
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "dct3d_golden.npz"))
    for i in range(g["blob_hdr"].shape[0]):  # step tables the reference decoder built for its own blobs
        t = workloads.step_table_qstep(g["blob_hdr"][i, 2])
        assert np.array_equal(t.view(np.uint32), g["blob_steps"][i].view(np.uint32)), i
