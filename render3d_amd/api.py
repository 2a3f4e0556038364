"""Python host API over the C ABI (include/r3d_hip.h).

Device buffers are torch tensors on the GPU (torch is used for allocation and
streams only); every computation runs in libr3d_hip.so. Mirrors the reference's
render backend calls (src/render/render.h) and dct3d's decode
(tools/3ddct/dct3d.h:80-91), batched.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import (BlobBatch, BlobParseOut, ChunkBatch, DecodeOpts, FrameParams, Image, R3DError, RaycastOpts,
                   RaycastTuning, Stats,
                   VolumeDesc)

try:
    import torch
except ImportError:  # pragma: no cover
    torch = None


def _ptr(t):
    return None if t is None else C.c_void_p(t.data_ptr())


def _stream(stream):
    """A HIP stream handle: an int, a torch stream, or None for torch's current stream."""
    if stream is not None:
        return C.c_void_p(getattr(stream, "cuda_stream", stream))
    if torch is not None and torch.cuda.is_available():
        return C.c_void_p(torch.cuda.current_stream().cuda_stream)
    return None


def headers_tensor(vmin, vspan, dcq, table=None, device="cuda"):
    """r3d_chunk_header[n] as an int32 [n, 4] tensor (floats bit-cast)."""
    n = len(vmin)
    h = np.zeros((n, 4), np.int32)
    h[:, 0] = np.asarray(vmin, np.float32).view(np.int32)
    h[:, 1] = np.asarray(vspan, np.float32).view(np.int32)
    h[:, 2] = np.asarray(dcq, np.int32)
    h[:, 3] = 0 if table is None else np.asarray(table, np.uint32).view(np.int32)
    return torch.from_numpy(h).to(device)


def headers_from_tensor(h):
    a = h.cpu().numpy()
    return (a[:, 0].view(np.float32).copy(), a[:, 1].view(np.float32).copy(), a[:, 2].copy(),
            a[:, 3].view(np.uint32).copy())


class Context:
    """r3d_hip_ctx: one device, one host thread at a time."""

    def __init__(self, device=0):
        self.lib = _lib.load()
        h = C.c_void_p()
        rc = self.lib.r3d_hip_init(C.byref(h), device)
        if rc != 0:
            raise R3DError(rc, f"r3d_hip_init(device={device}) failed")
        self.h = h
        self.device = device

    def close(self):
        if self.h:
            self.lib.r3d_hip_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def check(self, rc):
        if rc != 0:
            raise R3DError(rc, self.lib.r3d_hip_last_error(self.h).decode())
        return rc

    def sync(self, stream=None):
        self.check(self.lib.r3d_hip_sync(self.h, _stream(stream)))

    def dct_basis(self):
        out = np.zeros(256, np.float32)
        self.check(self.lib.r3d_hip_dct_basis(self.h, out.ctypes.data_as(C.c_void_p)))
        return out

    # ---- decode -----------------------------------------------------------
    def make_batch(self, *, headers, step_tables, out, levels=None, pairs=None, pair_offsets=None,
                   out_offsets=None, stride_y=0, stride_z=0, corr_offsets=None, corr_pos=None,
                   corr_delta=None, status=None, nchunks=None, pair_counts=None, corr_counts=None):
        b = ChunkBatch()
        b.nchunks = int(headers.shape[0]) if nchunks is None else int(nchunks)
        if levels is not None:
            b.coef_format = _lib.R3D_COEF_DENSE_I16
            b.levels = levels.data_ptr()
        else:
            b.coef_format = _lib.R3D_COEF_SPARSE
            b.pairs = pairs.data_ptr()
            b.pair_offsets = pair_offsets.data_ptr()
            b.npairs = int(pairs.numel())
        b.headers = headers.data_ptr()
        b.step_tables = step_tables.data_ptr()
        b.ntables = int(step_tables.numel() // 46)
        if corr_offsets is not None:
            b.corr_offsets = corr_offsets.data_ptr()
            b.corr_pos = corr_pos.data_ptr()
            b.corr_delta = corr_delta.data_ptr()
            b.ncorr = int(corr_pos.numel())
        b.out = out if isinstance(out, int) else out.data_ptr()
        b.out_offsets = None if out_offsets is None else out_offsets.data_ptr()
        b.out_stride_y = stride_y
        b.out_stride_z = stride_z
        b.status = None if status is None else status.data_ptr()
        b.pair_counts = None if pair_counts is None else pair_counts.data_ptr()
        b.corr_counts = None if corr_counts is None else corr_counts.data_ptr()
        return b

    def decode(self, batch, variant=_lib.R3D_DECODE_EXACT_F32, stream=None, flags=0):
        opts = DecodeOpts(variant, flags)
        self.check(self.lib.r3d_hip_decode_chunks(self.h, C.byref(batch), C.byref(opts), _stream(stream)))

    def decode_chunks(self, *, headers, step_tables, levels=None, pairs=None, pair_offsets=None,
                      corr_offsets=None, corr_pos=None, corr_delta=None, status=None,
                      variant=_lib.R3D_DECODE_EXACT_F32, stream=None, f32_out=False):
        """Decode into a fresh flat [n, 4096] uint8 tensor, or with f32_out the
        float32 value before rounding (R3D_DECODE_F32_OUT)."""
        n = int(headers.shape[0])
        out = torch.empty((n, 4096), dtype=torch.float32 if f32_out else torch.uint8, device=headers.device)
        b = self.make_batch(headers=headers, step_tables=step_tables, out=out, levels=levels, pairs=pairs,
                            pair_offsets=pair_offsets, corr_offsets=corr_offsets, corr_pos=corr_pos,
                            corr_delta=corr_delta, status=status)
        self.decode(b, variant, stream, _lib.R3D_DECODE_F32_OUT if f32_out else 0)
        return out

    # ---- 3ddct blobs -------------------------------------------------------
    def encode_blobs(self, levels, headers, qstep, *, dtype=_lib.R3D_DT_U8, slot_bytes=4096 * 8 + 256, stream=None):
        """Entropy-code dense levels [n, 4096] int16 with headers [n, 4] (vmin,
        vspan, dcq) and qstep (a float or a [n] float32 tensor) into 3ddct
        blobs. Returns (bytes uint8 [n * slot_bytes], offsets int64 [n],
        lengths int32 [n]); blob c starts at c * slot_bytes."""
        n = int(levels.shape[0])
        dev = levels.device
        qs = qstep if isinstance(qstep, torch.Tensor) else torch.full((n,), float(qstep), dtype=torch.float32,
                                                                        device=dev)
        out = torch.empty(n * slot_bytes, dtype=torch.uint8, device=dev)
        lengths = torch.empty(n, dtype=torch.int32, device=dev)
        self.check(self.lib.r3d_hip_encode_blobs(self.h, n, _ptr(levels), _ptr(headers), _ptr(qs), dtype, _ptr(out),
                                                 slot_bytes, _ptr(lengths), _stream(stream)))
        offsets = torch.arange(n, dtype=torch.int64, device=dev) * slot_bytes
        return out, offsets, lengths

    def parse_blobs(self, data, offsets, lengths, *, max_pairs=4096, dtype=_lib.R3D_DT_U8, stream=None,
                    noalloc=False):
        """r3d_hip_parse_blobs (or, with noalloc, r3d_hip_parse_blobs_noalloc):
        blobs -> the arrays of a sparse decode batch (dict of tensors, plus
        'status')."""
        n = int(offsets.shape[0])
        dev = data.device
        o = {"headers": torch.empty((n, 4), dtype=torch.int32, device=dev),
             "step_tables": torch.empty(n * 46, dtype=torch.float32, device=dev),
             "pair_offsets": torch.empty(n + 1, dtype=torch.int32, device=dev),
             "pair_counts": torch.empty(n, dtype=torch.int32, device=dev),
             "pairs": torch.empty(n * max_pairs, dtype=torch.int32, device=dev),
             "corr_offsets": torch.empty(n + 1, dtype=torch.int32, device=dev),
             "corr_counts": torch.empty(n, dtype=torch.int32, device=dev),
             "corr_pos": torch.empty(n * max_pairs, dtype=torch.int16, device=dev),
             "corr_delta": torch.empty(n * max_pairs, dtype=torch.float32, device=dev),
             "status": torch.empty(n, dtype=torch.int32, device=dev)}
        bb = BlobBatch(n, dtype, max_pairs, 0, _ptr(data), _ptr(offsets), _ptr(lengths))
        po = BlobParseOut(*[_ptr(o[k]) for k in ("headers", "step_tables", "pair_offsets", "pair_counts", "pairs",
                                                  "corr_offsets", "corr_counts", "corr_pos", "corr_delta",
                                                  "status")])
        fn = self.lib.r3d_hip_parse_blobs_noalloc if noalloc else self.lib.r3d_hip_parse_blobs
        self.check(fn(self.h, C.byref(bb), C.byref(po), _stream(stream)))
        return o

    def parsed_batch(self, parsed, out, **kw):
        """A sparse r3d_chunk_batch over parse_blobs' output."""
        return self.make_batch(headers=parsed["headers"], step_tables=parsed["step_tables"], out=out,
                               pairs=parsed["pairs"], pair_offsets=parsed["pair_offsets"],
                               pair_counts=parsed["pair_counts"], corr_offsets=parsed["corr_offsets"],
                               corr_pos=parsed["corr_pos"], corr_delta=parsed["corr_delta"],
                               corr_counts=parsed["corr_counts"], **kw)

    def decode_blobs(self, data, offsets, lengths, *, max_pairs=4096, variant=_lib.R3D_DECODE_EXACT_F32,
                     stream=None):
        """3ddct blobs -> u8 [n, 4096] (parse, then decode) and the parse status."""
        parsed = self.parse_blobs(data, offsets, lengths, max_pairs=max_pairs, stream=stream)
        n = int(offsets.shape[0])
        out = torch.empty((n, 4096), dtype=torch.uint8, device=data.device)
        self.decode(self.parsed_batch(parsed, out), variant, stream)
        return out, parsed["status"]

    # ---- zarr v3 shards ---------------------------------------------------------
    def crc32c(self, data, stream=None):
        """r3d_hip_crc32c of a uint8 device tensor -> int (synchronises)."""
        out = torch.zeros(1, dtype=torch.int32, device=data.device)
        self.check(self.lib.r3d_hip_crc32c(self.h, _ptr(data), int(data.numel()), _ptr(out), _stream(stream)))
        return int(out.cpu().numpy().view(np.uint32)[0])

    @staticmethod
    def _region(dims, region):
        nz, ny, nx = (int(v) for v in dims)
        z0, y0, x0, dz, dy, dx = (int(v) for v in region)
        return _lib.ShardRegion(nz, ny, nx, z0, y0, x0, dz, dy, dx, 0)

    def shard_region_chunks(self, dims, region):
        r = self._region(dims, region)
        return int(self.lib.r3d_hip_shard_region_chunks(C.byref(r)))

    def shard_decode_region(self, shards, dims, region, *, out=None, counts=None, status=True, stream=None):
        """r3d_hip_shard_decode_region. shards: [(uint8 device tensor: the
        whole shard file, sz, sy, sx)] (unlisted shards are absent); dims
        (nz, ny, nx); region (z0, y0, x0, dz, dy, dx). Returns (dst uint8
        [dz, dy, dx], counts int32 [4]: decoded, missing, corrupt, beyond
        int16, per-chunk status int32 or None); nothing is synchronised."""
        r = self._region(dims, region)
        dev = f"cuda:{self.device}"
        if out is None:
            out = torch.empty((r.dz, r.dy, r.dx), dtype=torch.uint8, device=dev)
        if counts is None:
            counts = torch.zeros(4, dtype=torch.int32, device=dev)
        st = None
        if status:
            st = torch.empty(max(1, int(self.lib.r3d_hip_shard_region_chunks(C.byref(r)))), dtype=torch.int32,
                             device=dev)
        arr = (_lib.ShardSrc * max(1, len(shards)))()
        for i, (t, sz, sy, sx) in enumerate(shards):
            arr[i] = _lib.ShardSrc(t.data_ptr(), int(t.numel()), int(sz), int(sy), int(sx), 0)
        self.check(self.lib.r3d_hip_shard_decode_region(self.h, C.cast(arr, C.c_void_p), len(shards), C.byref(r),
                                                        _ptr(out), _ptr(st), _ptr(counts), _stream(stream)))
        return out, counts, st

    # ---- deblocking ----------------------------------------------------------
    def deblock(self, vol, step, strength=1.0, stream=None):
        """r3d_hip_deblock_u8 in place on a uint8 device tensor [nz, ny, nx]."""
        assert vol.dtype == torch.uint8 and vol.dim() == 3 and vol.is_contiguous()
        dims = (C.c_uint64 * 3)(*vol.shape)
        self.check(self.lib.r3d_hip_deblock_u8(self.h, _ptr(vol), dims, float(step), float(strength),
                                               _stream(stream)))
        return vol

    def deblock_chunks(self, chunks, neighbors, step, strength=1.0, out=None, stream=None):
        """r3d_hip_deblock_chunks_u8: chunks uint8 [n, 4096], neighbors int32
        [n, 6] (xlo, xhi, ylo, yhi, zlo, zhi; -1 for none) -> uint8 [n, 4096]."""
        n = int(chunks.shape[0])
        if out is None:
            out = torch.empty_like(chunks)
        self.check(self.lib.r3d_hip_deblock_chunks_u8(self.h, n, _ptr(chunks), _ptr(neighbors), float(step),
                                                      float(strength), _ptr(out), _stream(stream)))
        return out

    # ---- workloads ----------------------------------------------------------
    def synth_sheets(self, nx, ny, nz, nsheets, seed, stream=None):
        out = torch.empty(nx * ny * nz, dtype=torch.uint8, device=f"cuda:{self.device}")
        self.check(self.lib.r3d_hip_synth_sheets(self.h, _ptr(out), nx, ny, nz, nsheets, seed, _stream(stream)))
        return out

    def encode_chunks(self, vol, nx, ny, nz, step_table, stream=None):
        n = (nx // 16) * (ny // 16) * (nz // 16)
        levels = torch.empty((n, 4096), dtype=torch.int16, device=vol.device)
        hdr = torch.empty((n, 4), dtype=torch.int32, device=vol.device)
        self.check(self.lib.r3d_hip_encode_chunks(self.h, _ptr(vol), nx, ny, nz, _ptr(step_table), _ptr(levels),
                                                  _ptr(hdr), _stream(stream)))
        return levels, hdr

    # ---- transfer / volumes / raycast ------------------------------------
    def set_transfer(self, lut_u8):
        lut = np.ascontiguousarray(lut_u8, np.uint8).reshape(256, 4)
        self.check(self.lib.r3d_hip_set_transfer(self.h, lut.ctypes.data_as(C.c_void_p)))

    def set_transfer_f32(self, lut):
        lut = np.ascontiguousarray(lut, np.float32).reshape(256, 4)
        self.check(self.lib.r3d_hip_set_transfer_f32(self.h, lut.ctypes.data_as(C.c_void_p)))

    def volume(self, nx, ny, nz, nlevels=0):
        return Volume(self, nx, ny, nz, nlevels)

    def raycast(self, vol, views, width, height, *, alpha_cutoff=0.98, rgba=True, rgba8=False, tf=None,
                counters=None, px_steps=False, stats=False, out=None, stream=None, measure=False, touched=None,
                pane_cache=False, out8=None):
        """Render views into a [height, width, 4] float32 image (and/or RGBA8).
        measure=True runs the measurement instances (R3D_RAYCAST_MEASURE):
        stats and counters then include the blended-sample counts, and
        touched (an int32 device bitmap of vol.brick_count bits, zeroed)
        receives the bricks the footprints read. pane_cache=True
        (R3D_RAYCAST_PANE_CACHE) draws only the views whose inputs changed
        since they were last drawn into the same out / out8; stats then
        count the drawn views (panes_drawn)."""
        return self._raycast(self.lib.r3d_hip_raycast, vol.h, views, width, height, alpha_cutoff, rgba, rgba8, tf,
                             counters, px_steps, stats, out, stream, measure, touched, pane_cache, out8)

    def invalidate_panes(self):
        """r3d_hip_invalidate_panes: every pane-cached view redraws next time."""
        self.lib.r3d_hip_invalidate_panes(self.h)

    def _raycast(self, fn, handle, views, width, height, alpha_cutoff, rgba, rgba8, tf, counters, px_steps, stats,
                 out, stream, measure=False, touched=None, pane_cache=False, out8=None):
        if isinstance(views, FrameParams):
            views = [views]
        arr = (FrameParams * len(views))(*views)
        dev = f"cuda:{self.device}"
        img = out if out is not None else (torch.empty((height, width, 4), dtype=torch.float32, device=dev)
                                          if rgba else None)
        img8 = out8 if out8 is not None else (torch.zeros((height, width, 4), dtype=torch.uint8, device=dev)
                                              if rgba8 else None)
        steps = torch.zeros((height, width), dtype=torch.int32, device=dev) if px_steps else None
        im = Image(img.data_ptr() if img is not None else None, img8.data_ptr() if img8 is not None else None,
                   width, height)
        flags = (_lib.RAYCAST_MEASURE if measure else 0) | (_lib.RAYCAST_PANE_CACHE if pane_cache else 0)
        opts = RaycastOpts(alpha_cutoff, flags, None if tf is None else tf.data_ptr(),
                           None if counters is None else counters.data_ptr(),
                           None if steps is None else steps.data_ptr(),
                           None if touched is None else touched.data_ptr())
        st = Stats()
        self.check(fn(self.h, handle, arr, len(views), C.byref(im), C.byref(opts), C.byref(st) if stats else None,
                      _stream(stream)))
        res = {"rgba": img, "rgba8": img8, "px_steps": steps}
        if stats:
            res["stats"] = {"samples": st.samples, "skips": st.skips, "shaded": st.shaded, "rays": st.rays,
                            "gpu_raycast_ns": st.gpu_raycast_ns, "panes_drawn": st.panes_drawn}
            if measure:
                res["stats"].update(blended=st.blended, blended_shaded=st.blended_shaded)
        return res

    def set_quality(self, quality):
        """r3d_set_quality: QUALITY_FULL (six gradient taps) or QUALITY_FAST
        (four tetrahedral taps) for later renders and plans."""
        _lib.load().r3d_hip_set_quality(self.h, quality)

    def set_raycast_tuning(self, tile_w, tile_h, blocks_per_cu=0):
        """Launch shape of raycast() (r3d_hip_set_raycast_tuning)."""
        self.check(self.lib.r3d_hip_set_raycast_tuning(self.h, C.byref(RaycastTuning(tile_w, tile_h,
                                                                                     blocks_per_cu, 0.0))))

    def raycast_plan(self, vol, views, width, height, *, alpha_cutoff=0.98, rgba=True, rgba8=False, tf=None,
                     counters=None, px_steps=False):
        """A capture-safe RaycastPlan over fixed views; owns its output tensors."""
        return RaycastPlan(self, vol, views, width, height, alpha_cutoff=alpha_cutoff, rgba=rgba, rgba8=rgba8,
                           tf=tf, counters=counters, px_steps=px_steps)


class RaycastPlan:
    """r3d_hip_raycast_plan: views and output fixed at creation; launch() is
    kernel launches only, so it can be captured into a HIP graph."""

    def __init__(self, ctx, vol, views, width, height, *, alpha_cutoff, rgba, rgba8, tf, counters, px_steps):
        if isinstance(views, FrameParams):
            views = [views]
        self.ctx = ctx
        self.vol = vol
        dev = f"cuda:{ctx.device}"
        self.rgba = torch.zeros((height, width, 4), dtype=torch.float32, device=dev) if rgba else None
        self.rgba8 = torch.zeros((height, width, 4), dtype=torch.uint8, device=dev) if rgba8 else None
        self.px_steps = torch.zeros((height, width), dtype=torch.int32, device=dev) if px_steps else None
        self.tf = tf
        self.counters = counters
        arr = (FrameParams * len(views))(*views)
        im = Image(None if self.rgba is None else self.rgba.data_ptr(),
                   None if self.rgba8 is None else self.rgba8.data_ptr(), width, height)
        opts = RaycastOpts(alpha_cutoff, 0, None if tf is None else tf.data_ptr(),
                           None if counters is None else counters.data_ptr(),
                           None if self.px_steps is None else self.px_steps.data_ptr())
        h = C.c_void_p()
        ctx.check(ctx.lib.r3d_hip_raycast_plan_create(ctx.h, vol.h, arr, len(views), C.byref(im), C.byref(opts),
                                                      C.byref(h)))
        self.h = h

    def launch(self, stream=None):
        self.ctx.check(self.ctx.lib.r3d_hip_raycast_plan_launch(self.h, _stream(stream)))

    def tune(self, stream=None):
        """Run the autotuner; returns every measured shape as a dict (the fastest is kept)."""
        res = (RaycastTuning * 64)()
        n = C.c_uint32()
        self.ctx.check(self.ctx.lib.r3d_hip_raycast_plan_tune(self.h, _stream(stream), res, 64, C.byref(n)))
        return [{"tile": (r.tile_w, r.tile_h), "blocks_per_cu": r.blocks_per_cu, "ms": r.ms} for r in res[:n.value]]

    @property
    def tuning(self):
        t = RaycastTuning()
        self.ctx.check(self.ctx.lib.r3d_hip_raycast_plan_get_tuning(self.h, C.byref(t)))
        return {"tile": (t.tile_w, t.tile_h), "blocks_per_cu": t.blocks_per_cu, "ms": t.ms}

    def set_tuning(self, tile_w, tile_h, blocks_per_cu=0):
        self.ctx.check(self.ctx.lib.r3d_hip_raycast_plan_set_tuning(
            self.h, C.byref(RaycastTuning(tile_w, tile_h, blocks_per_cu, 0.0))))

    def close(self):
        if self.h:
            self.ctx.lib.r3d_hip_raycast_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Volume:
    """r3d_hip_volume: bricked pyramid + occupancy in HBM."""

    def __init__(self, ctx, nx, ny, nz, nlevels=0):
        self.ctx = ctx
        d = VolumeDesc(nx, ny, nz, 128, 1.0)
        h = C.c_void_p()
        ctx.check(ctx.lib.r3d_hip_volume_create(ctx.h, C.byref(d), nlevels, C.byref(h)))
        self.h = h
        self.dims_xyz = (nx, ny, nz)
        nl = C.c_uint32()
        dims = (C.c_uint32 * 48)()
        od = (C.c_uint32 * 3)()
        ctx.check(ctx.lib.r3d_hip_volume_info(h, C.byref(nl), dims, od))
        self.nlevels = nl.value
        self.level_dims = [tuple(dims[l * 3:l * 3 + 3]) for l in range(self.nlevels)]
        self.odim = tuple(od)

    def close(self):
        if self.h:
            self.ctx.lib.r3d_hip_volume_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, voxels, stream=None):
        if torch is not None and isinstance(voxels, torch.Tensor) and voxels.is_cuda:
            self.ctx.check(self.ctx.lib.r3d_hip_volume_upload(self.ctx.h, self.h, _ptr(voxels), _lib.R3D_SRC_DEVICE,
                                                              _stream(stream)))
        else:
            v = np.ascontiguousarray(voxels, np.uint8)
            self.ctx.check(self.ctx.lib.r3d_hip_volume_upload(self.ctx.h, self.h, v.ctypes.data_as(C.c_void_p),
                                                              _lib.R3D_SRC_HOST, _stream(stream)))

    def chunk_targets(self, stream=None):
        nx, ny, nz = self.dims_xyz
        n = (nx // 16) * (ny // 16) * (nz // 16)
        offs = torch.empty(n, dtype=torch.int64, device=f"cuda:{self.ctx.device}")
        base = C.c_void_p()
        sy, sz = C.c_uint32(), C.c_uint32()
        self.ctx.check(self.ctx.lib.r3d_hip_volume_chunk_targets(self.ctx.h, self.h, _ptr(offs), C.byref(base),
                                                                 C.byref(sy), C.byref(sz), _stream(stream)))
        return offs, base.value, sy.value, sz.value

    def build(self, stream=None):
        """Aprons, pyramid and occupancy after level 0 was filled via chunk_targets()."""
        self.ctx.check(self.ctx.lib.r3d_hip_volume_build(self.ctx.h, self.h, _stream(stream)))

    def build_from_chunks(self, chunks, stream=None):
        """Whole volume from flat decoded chunks ([n, 4096] uint8, brick order)."""
        self.ctx.check(self.ctx.lib.r3d_hip_volume_build_from_chunks(self.ctx.h, self.h, _ptr(chunks),
                                                                     _stream(stream)))

    def decode(self, batch, variant=_lib.R3D_DECODE_EXACT_F32, stream=None):
        """Decode every chunk of the volume (brick order) and build it."""
        opts = DecodeOpts(variant, 0)
        self.ctx.check(self.ctx.lib.r3d_hip_volume_decode(self.ctx.h, self.h, C.byref(batch), C.byref(opts),
                                                          _stream(stream)))

    def decode_shards(self, shards, dims, origin=(0, 0, 0), counts=None, stream=None):
        """r3d_hip_volume_decode_shards: the volume's box at origin (z0, y0, x0)
        of a shard store of dims (nz, ny, nx), decoded and built. shards as
        Context.shard_decode_region; returns the counts (int32 [4])."""
        ctx = self.ctx
        nx, ny, nz = self.dims_xyz
        r = ctx._region(dims, tuple(origin) + (nz, ny, nx))
        if counts is None:
            counts = torch.zeros(4, dtype=torch.int32, device=f"cuda:{ctx.device}")
        arr = (_lib.ShardSrc * max(1, len(shards)))()
        for i, (t, sz, sy, sx) in enumerate(shards):
            arr[i] = _lib.ShardSrc(t.data_ptr(), int(t.numel()), int(sz), int(sy), int(sx), 0)
        ctx.check(ctx.lib.r3d_hip_volume_decode_shards(ctx.h, self.h, C.cast(arr, C.c_void_p), len(shards),
                                                       C.byref(r), None, _ptr(counts), _stream(stream)))
        return counts

    @property
    def brick_count(self):
        """16^3 bricks over all levels (r3d_hip_volume_brick_count)."""
        return int(self.ctx.lib.r3d_hip_volume_brick_count(self.h))

    def read_level(self, level, stream=None):
        nx, ny, nz = self.level_dims[level]
        out = np.empty(nx * ny * nz, np.uint8)
        self.ctx.check(self.ctx.lib.r3d_hip_volume_read_level(self.ctx.h, self.h, level,
                                                              out.ctypes.data_as(C.c_void_p), _lib.R3D_SRC_HOST,
                                                              _stream(stream)))
        return out

    def read_occupancy(self, stream=None):
        ox, oy, oz = self.odim
        out = np.empty(ox * oy * oz, np.uint8)
        self.ctx.check(self.ctx.lib.r3d_hip_volume_read_occupancy(self.ctx.h, self.h,
                                                                  out.ctypes.data_as(C.c_void_p), _lib.R3D_SRC_HOST,
                                                                  _stream(stream)))
        return out


class Bricks:
    """r3d_hip_bricks: a volume as a pyramid of 128^3 bricks in the slots of
    an atlas, found through the reference's page table (vkbackend.c:4185-4205),
    rendered in bricks-LOD mode (raycast.slang R3D_MODE 3 with BR_LOD_FLAG)."""

    def __init__(self, ctx, nx, ny, nz, nlev, abpa):
        self.ctx = ctx
        d = _lib.BricksDesc(nx, ny, nz, nlev, abpa, 0)
        h = C.c_void_p()
        ctx.check(ctx.lib.r3d_hip_bricks_create(ctx.h, C.byref(d), C.byref(h)))
        self.h = h
        self.dims_xyz, self.nlev, self.abpa = (nx, ny, nz), nlev, abpa

    def close(self):
        if self.h:
            self.ctx.lib.r3d_hip_bricks_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def brick_mode(self):
        """The brick_mode a view of this atlas carries."""
        return self.abpa | _lib.BR_LOD_FLAG

    def tables(self):
        """(page table device address, its words, atlas device address, slots)."""
        page, atlas = C.c_void_p(), C.c_void_p()
        words, nslots = C.c_uint32(), C.c_uint32()
        self.ctx.check(self.ctx.lib.r3d_hip_bricks_tables(self.h, C.byref(page), C.byref(words), C.byref(atlas),
                                                          C.byref(nslots)))
        return page.value, words.value, atlas.value, nslots.value

    def build(self, vol, resident=None, stream=None):
        """Pyramid of vol (uint8 [nz, ny, nx] on the device); bricks whose
        resident flag (one per page entry, page order) is 0 stay absent."""
        nx, ny, nz = self.dims_xyz
        if not (isinstance(vol, torch.Tensor) and vol.is_cuda and vol.dtype == torch.uint8 and vol.is_contiguous()
                and vol.numel() == nx * ny * nz):
            raise ValueError(f"bricks build: need a contiguous uint8 device tensor of {nz}x{ny}x{nx}")
        res = None
        if resident is not None:
            res = np.ascontiguousarray(resident, np.uint8)
            if res.size != self.tables()[1] - 64:
                raise ValueError("bricks build: one resident flag per page entry")
        self.ctx.check(self.ctx.lib.r3d_hip_bricks_build(self.ctx.h, self.h, _ptr(vol),
                                                         None if res is None else res.ctypes.data_as(C.c_void_p),
                                                         _stream(stream)))

    def build_from_shards(self, levels, resident=None, stream=None):
        """r3d_hip_bricks_build_from_shards: levels = one shard list per level,
        [(uint8 device tensor: the whole shard file, sz, sy, sx)], the store's
        volume being that level's dims; streaming then decodes bricks from
        them. Bricks whose resident flag is set are decoded now."""
        if len(levels) != self.nlev:
            raise ValueError(f"bricks build_from_shards: {len(levels)} level stores for {self.nlev} levels")
        self._shard_keep = []
        lv = (_lib.BricksLevelSrc * self.nlev)()
        for l, shards in enumerate(levels):
            arr = (_lib.ShardSrc * max(1, len(shards)))()
            for i, (t, sz, sy, sx) in enumerate(shards):
                arr[i] = _lib.ShardSrc(t.data_ptr(), int(t.numel()), int(sz), int(sy), int(sx), 0)
            self._shard_keep.append((arr, list(shards)))
            lv[l] = _lib.BricksLevelSrc(C.cast(arr, C.c_void_p), len(shards), 0)
        res = None
        if resident is not None:
            res = np.ascontiguousarray(resident, np.uint8)
            if res.size != self.tables()[1] - 64:
                raise ValueError("bricks build_from_shards: one resident flag per page entry")
        self.ctx.check(self.ctx.lib.r3d_hip_bricks_build_from_shards(
            self.ctx.h, self.h, lv, self.nlev, None if res is None else res.ctypes.data_as(C.c_void_p),
            _stream(stream)))

    def read_page(self):
        """The page table as uint32 on the host (synchronous)."""
        page, words, _, _ = self.tables()
        out = np.empty(words, np.uint32)
        _hip_copy(out.ctypes.data, page, out.nbytes)
        return out

    def read_slot(self, slot):
        """Slot `slot` of the atlas as uint8 [128, 128, 128] (z, y, x) on the host."""
        return self.read_slots(slot, 1)[0]

    def read_slots(self, first, n):
        """Slots first..first+n-1 as uint8 [n, 128, 128, 128] on the host (synchronous)."""
        _, _, atlas, nslots = self.tables()
        if not (0 <= first and n >= 0 and first + n <= nslots):
            raise IndexError((first, n))
        out = np.empty((n, 128, 128, 128), np.uint8)
        _hip_copy(out.ctypes.data, atlas + first * 128 ** 3, out.nbytes)
        return out

    def write_page(self, page):
        """Replace the whole page table (uint32, the table's length) from the host (synchronous)."""
        dst, words, _, _ = self.tables()
        page = np.ascontiguousarray(page, np.uint32)
        if page.size != words:
            raise ValueError(f"bricks write_page: {page.size} words for a table of {words}")
        _hip_copy(dst, page.ctypes.data, page.nbytes)

    def write_slots(self, first, voxels):
        """Fill slots first.. from host uint8 data of whole 128^3 slots (synchronous)."""
        _, _, atlas, nslots = self.tables()
        v = np.ascontiguousarray(voxels, np.uint8).reshape(-1)
        n = v.size // (128 ** 3)
        if v.size % (128 ** 3) or first < 0 or first + n > nslots:
            raise ValueError("bricks write_slots: whole slots inside the atlas")
        _hip_copy(atlas + first * 128 ** 3, v.ctypes.data, v.nbytes)

    # ---- streaming residency (r3d_hip_bricks_stream_*) --------------------------
    def stream_begin(self):
        """r3d_bricks_stream_begin: opens a collect pass (True) or not (False)."""
        return bool(self.ctx.lib.r3d_hip_bricks_stream_begin(self.h))

    def stream_box(self, lo, hi, pixel_cone, gate):
        f3 = C.c_float * 3
        self.ctx.lib.r3d_hip_bricks_stream_box(self.h, f3(*lo), f3(*hi), pixel_cone, gate)

    def stream_point(self, p, level, gate):
        self.ctx.lib.r3d_hip_bricks_stream_point(self.h, (C.c_float * 3)(*p), level, gate)

    def stream_submit(self, budget, stream=None):
        self.ctx.check(self.ctx.lib.r3d_hip_bricks_stream_submit(self.ctx.h, self.h, budget, _stream(stream)))

    def stream(self, eye, fwd, half_tan, pixel_cone, gate, budget, slice_z0=0, slice_depth=0, stream=None):
        """r3d_bricks_stream: begin + the view-cone collect + submit."""
        f3 = C.c_float * 3
        self.ctx.check(self.ctx.lib.r3d_hip_bricks_stream(self.ctx.h, self.h, f3(*eye), f3(*fwd), half_tan,
                                                          pixel_cone, slice_z0, slice_depth, gate, budget,
                                                          _stream(stream)))

    def stream_stats(self):
        st = _lib.BricksStreamStats()
        self.ctx.check(self.ctx.lib.r3d_hip_bricks_stream_stats(self.h, C.byref(st)))
        return {"nb": st.nb, "hot": st.hot, "hot_cap": st.hot_cap, "inflight": st.inflight,
                "evicted_last": st.evicted_last, "nlevels": st.nlevels, "lod_wanted": list(st.lod_wanted),
                "lod_requests": list(st.lod_requests), "decoded": st.decoded, "evicted": st.evicted,
                "jobs": st.jobs}

    def raycast(self, views, width, height, *, alpha_cutoff=0.98, rgba=True, rgba8=False, tf=None, counters=None,
                px_steps=False, stats=False, out=None, stream=None, pane_cache=False, out8=None):
        """Render bricks-LOD views (brick_mode = self.brick_mode); as Context.raycast."""
        return self.ctx._raycast(self.ctx.lib.r3d_hip_bricks_raycast, self.h, views, width, height, alpha_cutoff,
                                 rgba, rgba8, tf, counters, px_steps, stats, out, stream, pane_cache=pane_cache,
                                 out8=out8)


_hip = None


def _hip_copy(dst, src, nbytes):
    """hipMemcpy(dst, src, nbytes, hipMemcpyDefault) from the HIP runtime the library uses."""
    global _hip
    if _hip is None:
        _hip = C.CDLL("libamdhip64.so")
        _hip.hipMemcpy.restype = C.c_int
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    if _hip.hipMemcpy(dst, src, nbytes, 4) != 0:
        raise R3DError(-1, "hipMemcpy failed")


class ShardStore:
    """r3d_shard_store (src/shard/shardio.h:16-20) on the GPU: a directory of
    shard files <dir>/<Z>_<Y>_<X>.shard for a volume of nz x ny x nx voxels.
    A shard is read into HBM the first time a region needs it and kept."""

    def __init__(self, ctx, path, nz, ny, nx):
        self.ctx, self.path, self.dims = ctx, path, (int(nz), int(ny), int(nx))
        self.cache = {}

    def shard(self, sz, sy, sx):
        """The shard's bytes on the device, or None when its file is absent."""
        import os
        key = (int(sz), int(sy), int(sx))
        if key not in self.cache:
            f = os.path.join(self.path, f"{key[0]}_{key[1]}_{key[2]}.shard")
            self.cache[key] = (torch.from_numpy(np.fromfile(f, np.uint8)).to(f"cuda:{self.ctx.device}")
                               if os.path.exists(f) else None)
        return self.cache[key]

    def decode_region(self, z0, y0, x0, dz, dy, dx, stream=None):
        """r3d_shard_decode_region (shardio.c:269-306): returns (rc, dst uint8
        [dz, dy, dx] on the device). rc is 0, -1 when a shard or chunk was
        corrupt (those chunks are zero, as in the reference), or -2 when a
        chunk's levels exceed int16 (zero here; the reference decodes it)."""
        region = (z0, y0, x0, dz, dy, dx)
        nz, ny, nx = self.dims
        shards = []
        if z0 < nz and y0 < ny and x0 < nx:
            g = _lib.SHARD_GRID * 16
            ends = [min(a + d, n) - 1 for a, d, n in ((z0, dz, nz), (y0, dy, ny), (x0, dx, nx))]
            for sz in range(z0 // g, ends[0] // g + 1):
                for sy in range(y0 // g, ends[1] // g + 1):
                    for sx in range(x0 // g, ends[2] // g + 1):
                        t = self.shard(sz, sy, sx)
                        if t is not None:
                            shards.append((t, sz, sy, sx))
        dst, counts, _ = self.ctx.shard_decode_region(shards, self.dims, region, status=False, stream=stream)
        c = counts.cpu().numpy()
        return (-1 if c[2] else -2 if c[3] else 0), dst
