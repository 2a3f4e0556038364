"""Synthetic inputs for the batched path: sparse chunk batches with uneven
nonzero counts, and sets of views of mixed resolution packed into one image.

The reference has no generator for either; both follow the shapes of its
batched GPU decode (sparse (position, level) pairs per chunk,
src/vk/vkc5d.c:389-416) and of r3d_frame_views (views placed in one
framebuffer at their view_org, src/render/render.h:245-251).
"""
import numpy as np

from . import scene


def scan_order():
    """Ascending-L1-frequency scan of the 16^3 coefficient cube, ties in index
    order (the counting sort of tools/3ddct/dct3d.c:330-351)."""
    idx = np.arange(4096)
    l1 = idx // 256 + (idx // 16) % 16 + idx % 16
    return np.argsort(l1, kind="stable").astype(np.uint32)


def step_table_linear():
    """q(i, j, k) = 1 + 2 (i + j + k) as the 46-entry table indexed by i + j + k."""
    return (1.0 + 2.0 * np.arange(46)).astype(np.float32)


def step_table_qstep(qstep):
    """step[f] = qstep * powf(1 + f, 0.65f), f = 0..45: the table a 3ddct blob's
    qstep implies (tools/3ddct/dct3d.c:131-143). powf is the host libm's, the
    same values the blob parser's table holds."""
    import ctypes
    import ctypes.util
    libm = ctypes.CDLL(ctypes.util.find_library("m"))
    libm.powf.restype = ctypes.c_float
    libm.powf.argtypes = [ctypes.c_float, ctypes.c_float]
    q = np.float32(qstep)
    return np.array([q * np.float32(libm.powf(1.0 + f, 0.65)) for f in range(46)], np.float32)


def pack_shard(ctx, data, offsets, lengths, present=None):
    """A zarr v3 shard (the layout of tools/3ddct/tools/export/shard.c:55-97)
    from 262,144 blobs on the device (uint8 bytes, int64 offsets, int32
    lengths, in z-major chunk order): present chunks' blobs back to back, the
    index ((~0, ~0) where present is False) and its crc32c, as one uint8
    device tensor."""
    import torch
    n = 64 ** 3
    assert offsets.numel() == n and lengths.numel() == n
    dev = data.device
    keep = torch.ones(n, dtype=torch.bool, device=dev) if present is None else present.to(dev)
    l64 = torch.where(keep, lengths.long(), torch.zeros_like(offsets))
    out_off = torch.cumsum(l64, 0) - l64
    total = int(l64.sum())
    within = torch.arange(total, device=dev) - torch.repeat_interleave(out_off, l64)
    payload = data[torch.repeat_interleave(offsets, l64) + within]
    idx = torch.stack([out_off, l64], 1)
    idx[~keep] = -1  # (~0, ~0)
    ib = idx.contiguous().view(torch.uint8).reshape(-1)  # little-endian int64 pairs
    crc = ctx.crc32c(ib)
    tail = torch.tensor(np.array([crc], "<u4").view(np.uint8), device=dev)
    return torch.cat([payload, ib, tail])


def sparse_laplacian_batch(nchunks, seed, count_scale=100.0, value_scale=12.0, min_count=5, max_count=400):
    """A sparse chunk batch with uneven nonzero counts.

    Per chunk: count = clamp(round(|Laplace(0, count_scale)|), min_count,
    max_count); that many distinct positions drawn uniformly from the first
    4 * count entries of the scan order (so they sit in the low-L1 prefix, as
    an entropy-coded block's significant coefficients do); levels
    round(Laplace(0, value_scale)), zeros replaced by +-1, clipped to int16.
    Header: vmin in [0, 40), vspan in [120, 255), dc in [40, 160) in 1/16
    steps, table 0.

    Returns (offsets u32[n+1], pairs u32[nnz] as pos | level << 16,
    headers i32[n, 4] in r3d_chunk_header layout).
    """
    rng = np.random.default_rng(seed)
    counts = np.clip(np.rint(np.abs(rng.laplace(0.0, count_scale, nchunks))), min_count, max_count).astype(np.int64)
    scan = scan_order()
    offs = np.zeros(nchunks + 1, np.uint64)
    offs[1:] = np.cumsum(counts)
    pos = np.empty(int(offs[-1]), np.uint32)
    for c in range(nchunks):
        n = int(counts[c])
        pick = rng.choice(min(4096, 4 * n), n, replace=False)
        pos[int(offs[c]):int(offs[c]) + n] = np.sort(scan[pick])
    lv = np.rint(rng.laplace(0.0, value_scale, pos.size))
    lv[lv == 0] = np.where(rng.random(int((lv == 0).sum())) < 0.5, -1.0, 1.0)
    lv = np.clip(lv, -32768, 32767).astype(np.int16)
    pairs = pos | (lv.view(np.uint16).astype(np.uint32) << 16)
    hdr = np.zeros((nchunks, 4), np.int32)
    hdr[:, 0] = rng.uniform(0.0, 40.0, nchunks).astype(np.float32).view(np.int32)
    hdr[:, 1] = rng.uniform(120.0, 255.0, nchunks).astype(np.float32).view(np.int32)
    hdr[:, 2] = rng.integers(40 * 16, 160 * 16, nchunks)
    return offs.astype(np.uint32), pairs.astype(np.uint32), hdr


def select(offsets, pairs, headers, idx):
    """The chunks `idx` of a sparse batch as a batch of their own."""
    idx = np.asarray(idx, np.int64)
    cnt = np.diff(offsets.astype(np.int64))[idx]
    offs = np.zeros(idx.size + 1, np.uint32)
    offs[1:] = np.cumsum(cnt)
    sub = np.concatenate([pairs[offsets[i]:offsets[i + 1]] for i in idx]) if idx.size else np.zeros(0, np.uint32)
    return offs, sub.astype(np.uint32), np.ascontiguousarray(headers[idx])


def densify(offsets, pairs):
    """Sparse (offsets, pairs) -> dense int16 [n, 4096] levels."""
    n = offsets.size - 1
    out = np.zeros((n, 4096), np.int16)
    cid = np.repeat(np.arange(n), np.diff(offsets.astype(np.int64)))
    out[cid, pairs & 4095] = (pairs >> 16).astype(np.uint16).view(np.int16)
    return out


MIXED_VIEW_SIZES = ((16, 256, 256), (16, 512, 512), (16, 1280, 720), (16, 1920, 1080))


def mixed_views(sizes=MIXED_VIEW_SIZES, atlas_width=7680, step_voxels=1.0, gate=0.0, mode=None, seed=7):
    """Views of mixed resolution shelf-packed into one image of width
    atlas_width (rows left to right, each row as tall as its tallest view).
    Each view orbits the unit cube at its own yaw and pitch. Returns
    (views, width, height)."""
    rng = np.random.default_rng(seed)
    views = []
    x = y = row_h = 0
    i = 0
    total = sum(c for c, _, _ in sizes)
    for count, w, h in sizes:
        for _ in range(count):
            if x + w > atlas_width:
                x, y, row_h = 0, y + row_h, 0
            cam = scene.orbit_camera(360.0 * i / total, float(rng.uniform(-35.0, 35.0)),
                                     float(rng.uniform(1.6, 2.4)))
            kw = {} if mode is None else {"mode": mode}
            views.append(scene.frame_params(cam, w, h, step_voxels=step_voxels, gate=gate, view_org=(x, y),
                                            frame_index=i, **kw))
            x += w
            row_h = max(row_h, h)
            i += 1
    return views, atlas_width, y + row_h
