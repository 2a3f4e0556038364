"""ctypes binding of the C ABI in include/r3d_hip.h (libr3d_hip.so).

The library is the product: there is no Python or CPU fallback. If the shared
object is missing, loading raises, naming the build step.

torch (when importable) is imported first: it ships a HIP runtime with the same
SONAME (libamdhip64.so.7) as the one the library links, so importing it first
makes the process use one runtime and lets torch device tensors be passed to
the library as plain pointers.
"""
import ctypes as C
import os

try:  # one HIP runtime per process (see module docstring)
    import torch  # noqa: F401
except ImportError:  # pragma: no cover - torch is present in this image
    torch = None

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libr3d_hip.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "r3d_hip.h")

R3D_OK = 0
R3D_ERR_INVALID = -1
R3D_ERR_NOMEM = -2
R3D_ERR_DEVICE = -3
R3D_ERR_UNSUPPORTED = -4
R3D_ERR_RANGE = -5

R3D_COEF_DENSE_I16 = 0
R3D_COEF_SPARSE = 1
R3D_DECODE_EXACT_F32 = 0
R3D_DECODE_MFMA_F32 = 1
R3D_DECODE_MFMA_F16 = 2
R3D_DECODE_MFMA_F16X3 = 3
R3D_DECODE_F32_OUT = 1
R3D_SRC_HOST = 0
R3D_SRC_DEVICE = 1

MODE_FULL, MODE_MIP, MODE_DEPTH, MODE_HEATMAP, MODE_RAYDIR, MODE_FLAT = range(6)
VIEW_ORTHO = 1
VIEW_HALF = 32


class FrameParams(C.Structure):
    """r3d_frame_params, src/render/render_types.h:47-94 (248 bytes)."""

    _fields_ = [
        ("cam_origin", C.c_float * 3), ("step_voxels", C.c_float),
        ("cam_right", C.c_float * 3), ("density", C.c_float),
        ("cam_up", C.c_float * 3), ("lod_bias", C.c_float),
        ("cam_forward", C.c_float * 3), ("max_mip", C.c_float),
        ("viewport", C.c_uint32 * 2),
        ("mode", C.c_uint32),
        ("frame_index", C.c_uint32),
        ("threshold", C.c_float),
        ("slab_grid", C.c_uint32),
        ("slab_wz", C.c_uint32),
        ("slab_z0", C.c_float),
        ("slab_nx", C.c_float), ("slab_ny", C.c_float),
        ("slab_px", C.c_float), ("slab_py", C.c_float),
        ("slab_depth", C.c_uint32),
        ("clip_valid", C.c_uint32),
        ("clip_orig", (C.c_float * 2) * 6),
        ("vol_r0", C.c_float * 3), ("vol_tx", C.c_float),
        ("vol_r1", C.c_float * 3), ("vol_ty", C.c_float),
        ("vol_r2", C.c_float * 3), ("vol_tz", C.c_float),
        ("brick_mode", C.c_uint32),
        ("skip_gate", C.c_float),
        ("slab_x0", C.c_float), ("slab_y0", C.c_float),
        ("view_org", C.c_uint32),
        ("view_flags", C.c_uint32),
        ("overlay_gain", C.c_float),
        ("overlay_flags", C.c_uint32),
    ]


assert C.sizeof(FrameParams) == 248


class VolumeDesc(C.Structure):
    _fields_ = [("nx", C.c_uint32), ("ny", C.c_uint32), ("nz", C.c_uint32),
                ("brick_dim", C.c_uint32), ("voxel_um", C.c_float)]


class ChunkHeader(C.Structure):
    _fields_ = [("vmin", C.c_float), ("vspan", C.c_float), ("dcq", C.c_int32), ("table", C.c_uint32)]


class ChunkBatch(C.Structure):
    _fields_ = [
        ("nchunks", C.c_uint32), ("coef_format", C.c_uint32),
        ("levels", C.c_void_p), ("pair_offsets", C.c_void_p), ("pairs", C.c_void_p),
        ("npairs", C.c_uint64),
        ("headers", C.c_void_p), ("step_tables", C.c_void_p),
        ("ntables", C.c_uint32), ("reserved0", C.c_uint32),
        ("corr_offsets", C.c_void_p), ("corr_pos", C.c_void_p), ("corr_delta", C.c_void_p),
        ("ncorr", C.c_uint64),
        ("out", C.c_void_p), ("out_offsets", C.c_void_p),
        ("out_stride_y", C.c_uint32), ("out_stride_z", C.c_uint32),
        ("status", C.c_void_p),
        ("pair_counts", C.c_void_p), ("corr_counts", C.c_void_p),
    ]


class DecodeOpts(C.Structure):
    _fields_ = [("variant", C.c_uint32), ("flags", C.c_uint32)]


class Image(C.Structure):
    _fields_ = [("rgba", C.c_void_p), ("rgba8", C.c_void_p), ("width", C.c_uint32), ("height", C.c_uint32)]


class Stats(C.Structure):
    _fields_ = [("samples", C.c_uint64), ("skips", C.c_uint64), ("shaded", C.c_uint64), ("rays", C.c_uint64),
                ("blended", C.c_uint64), ("blended_shaded", C.c_uint64), ("gpu_raycast_ns", C.c_uint64),
                ("panes_drawn", C.c_uint32), ("reserved", C.c_uint32)]


RAYCAST_MEASURE = 1
RAYCAST_PANE_CACHE = 2


class RaycastOpts(C.Structure):
    _fields_ = [("alpha_cutoff", C.c_float), ("flags", C.c_uint32), ("tf", C.c_void_p),
                ("counters", C.c_void_p), ("px_steps", C.c_void_p), ("touched", C.c_void_p)]


class BlobBatch(C.Structure):
    _fields_ = [("nchunks", C.c_uint32), ("dtype", C.c_uint32), ("max_pairs", C.c_uint32), ("reserved0", C.c_uint32),
                ("bytes", C.c_void_p), ("offsets", C.c_void_p), ("lengths", C.c_void_p)]


class BlobParseOut(C.Structure):
    _fields_ = [("headers", C.c_void_p), ("step_tables", C.c_void_p), ("pair_offsets", C.c_void_p),
                ("pair_counts", C.c_void_p), ("pairs", C.c_void_p), ("corr_offsets", C.c_void_p),
                ("corr_counts", C.c_void_p), ("corr_pos", C.c_void_p), ("corr_delta", C.c_void_p),
                ("status", C.c_void_p)]


R3D_DT_U8 = 1
R3D_PARSE_NOALLOC_MAX_BLOBS = 262144
QUALITY_FULL, QUALITY_FAST = 0, 1  # render_types.h:35-38


class ShardSrc(C.Structure):
    _fields_ = [("bytes", C.c_void_p), ("nbytes", C.c_uint64), ("sz", C.c_uint32), ("sy", C.c_uint32),
                ("sx", C.c_uint32), ("reserved0", C.c_uint32)]


class BricksLevelSrc(C.Structure):
    _fields_ = [("shards", C.c_void_p), ("nshards", C.c_uint32), ("reserved0", C.c_uint32)]


class ShardRegion(C.Structure):
    _fields_ = [("nz", C.c_uint64), ("ny", C.c_uint64), ("nx", C.c_uint64), ("z0", C.c_uint64), ("y0", C.c_uint64),
                ("x0", C.c_uint64), ("dz", C.c_uint32), ("dy", C.c_uint32), ("dx", C.c_uint32),
                ("reserved0", C.c_uint32)]


class BricksDesc(C.Structure):
    _fields_ = [("nx", C.c_uint32), ("ny", C.c_uint32), ("nz", C.c_uint32), ("nlev", C.c_uint32),
                ("abpa", C.c_uint32), ("reserved0", C.c_uint32)]


BR_LOD_FLAG = 0x20000


class BricksStreamStats(C.Structure):
    _fields_ = [("nb", C.c_uint32), ("hot", C.c_uint32), ("hot_cap", C.c_uint32), ("inflight", C.c_uint32),
                ("evicted_last", C.c_uint32), ("nlevels", C.c_uint32), ("lod_wanted", C.c_uint32 * 8),
                ("lod_requests", C.c_uint64 * 8), ("decoded", C.c_uint64), ("evicted", C.c_uint64),
                ("jobs", C.c_uint64)]

SHARD_GRID = 64
SHARD_INDEX_BYTES = SHARD_GRID ** 3 * 16 + 4


class RaycastTuning(C.Structure):
    _fields_ = [("tile_w", C.c_uint32), ("tile_h", C.c_uint32), ("blocks_per_cu", C.c_uint32), ("ms", C.c_float)]


class DiffParams(C.Structure):
    _fields_ = [("tf", C.c_void_p), ("grid", C.c_void_p), ("grid_dims", C.c_uint32 * 3), ("alpha_cutoff", C.c_float),
                ("workspace", C.c_void_p), ("workspace_bytes", C.c_size_t)]


# name -> (restype, argtypes); also the export list the ABI test checks
P = C.c_void_p
SIGNATURES = {
    "r3d_hip_init": (C.c_int, [C.POINTER(P), C.c_int]),
    "r3d_hip_destroy": (None, [P]),
    "r3d_hip_last_error": (C.c_char_p, [P]),
    "r3d_hip_sync": (C.c_int, [P, P]),
    "r3d_hip_dct_basis": (C.c_int, [P, P]),
    "r3d_hip_decode_chunks": (C.c_int, [P, C.POINTER(ChunkBatch), C.POINTER(DecodeOpts), P]),
    "r3d_hip_volume_create": (C.c_int, [P, C.POINTER(VolumeDesc), C.c_uint32, C.POINTER(P)]),
    "r3d_hip_volume_destroy": (None, [P]),
    "r3d_hip_volume_info": (C.c_int, [P, P, P, P]),
    "r3d_hip_volume_upload": (C.c_int, [P, P, P, C.c_uint32, P]),
    "r3d_hip_volume_chunk_targets": (C.c_int, [P, P, P, C.POINTER(P), C.POINTER(C.c_uint32),
                                               C.POINTER(C.c_uint32), P]),
    "r3d_hip_volume_build": (C.c_int, [P, P, P]),
    "r3d_hip_volume_build_from_chunks": (C.c_int, [P, P, P, P]),
    "r3d_hip_volume_decode": (C.c_int, [P, P, C.POINTER(ChunkBatch), C.POINTER(DecodeOpts), P]),
    "r3d_hip_volume_read_level": (C.c_int, [P, P, C.c_uint32, P, C.c_uint32, P]),
    "r3d_hip_volume_brick_count": (C.c_uint64, [P]),
    "r3d_hip_volume_read_occupancy": (C.c_int, [P, P, P, C.c_uint32, P]),
    "r3d_hip_set_transfer": (C.c_int, [P, P]),
    "r3d_hip_set_transfer_f32": (C.c_int, [P, P]),
    "r3d_hip_raycast": (C.c_int, [P, P, C.POINTER(FrameParams), C.c_uint32, C.POINTER(Image),
                                  C.POINTER(RaycastOpts), C.POINTER(Stats), P]),
    "r3d_hip_set_raycast_tuning": (C.c_int, [P, C.POINTER(RaycastTuning)]),
    "r3d_hip_set_quality": (None, [P, C.c_uint32]),
    "r3d_hip_invalidate_panes": (None, [P]),
    "r3d_hip_raycast_plan_create": (C.c_int, [P, P, C.POINTER(FrameParams), C.c_uint32, C.POINTER(Image),
                                              C.POINTER(RaycastOpts), C.POINTER(P)]),
    "r3d_hip_raycast_plan_launch": (C.c_int, [P, P]),
    "r3d_hip_raycast_plan_tune": (C.c_int, [P, P, C.POINTER(RaycastTuning), C.c_uint32, C.POINTER(C.c_uint32)]),
    "r3d_hip_raycast_plan_set_tuning": (C.c_int, [P, C.POINTER(RaycastTuning)]),
    "r3d_hip_raycast_plan_get_tuning": (C.c_int, [P, C.POINTER(RaycastTuning)]),
    "r3d_hip_raycast_plan_destroy": (None, [P]),
    "r3d_hip_diff_workspace_bytes": (C.c_size_t, [C.POINTER(FrameParams), C.c_uint32]),
    "r3d_hip_raycast_diff": (C.c_int, [P, P, C.POINTER(FrameParams), C.c_uint32, C.POINTER(Image),
                                       C.POINTER(DiffParams), P]),
    "r3d_hip_raycast_adjoint": (C.c_int, [P, P, C.POINTER(FrameParams), C.c_uint32, P, C.c_uint32, C.c_uint32,
                                          P, P, C.POINTER(DiffParams), P]),
    "r3d_hip_synth_sheets": (C.c_int, [P, P, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                       C.c_uint64, P]),
    "r3d_hip_encode_chunks": (C.c_int, [P, P, C.c_uint32, C.c_uint32, C.c_uint32, P, P, P, P]),
    "r3d_hip_parse_blobs": (C.c_int, [P, C.POINTER(BlobBatch), C.POINTER(BlobParseOut), P]),
    "r3d_hip_parse_blobs_noalloc": (C.c_int, [P, C.POINTER(BlobBatch), C.POINTER(BlobParseOut), P]),
    "r3d_hip_encode_blobs": (C.c_int, [P, C.c_uint32, P, P, P, C.c_uint32, P, C.c_uint64, P, P]),
    "r3d_hip_shard_region_chunks": (C.c_uint64, [C.POINTER(ShardRegion)]),
    "r3d_hip_shard_decode_region": (C.c_int, [P, P, C.c_uint32, C.POINTER(ShardRegion), P, P, P, P]),
    "r3d_hip_crc32c": (C.c_int, [P, P, C.c_uint64, P, P]),
    "r3d_hip_deblock_u8": (C.c_int, [P, P, P, C.c_float, C.c_float, P]),
    "r3d_hip_bricks_create": (C.c_int, [P, C.POINTER(BricksDesc), C.POINTER(P)]),
    "r3d_hip_bricks_destroy": (None, [P]),
    "r3d_hip_bricks_build": (C.c_int, [P, P, P, P, P]),
    "r3d_hip_bricks_build_from_shards": (C.c_int, [P, P, P, C.c_uint32, P, P]),
    "r3d_hip_volume_decode_shards": (C.c_int, [P, P, P, C.c_uint32, P, P, P, P]),
    "r3d_hip_bricks_tables": (C.c_int, [P, C.POINTER(P), C.POINTER(C.c_uint32), C.POINTER(P), C.POINTER(C.c_uint32)]),
    "r3d_hip_bricks_raycast": (C.c_int, [P, P, C.POINTER(FrameParams), C.c_uint32, C.POINTER(Image),
                                         C.POINTER(RaycastOpts), C.POINTER(Stats), P]),
    "r3d_hip_deblock_chunks_u8": (C.c_int, [P, C.c_uint32, P, P, C.c_float, C.c_float, P, P]),
    "r3d_hip_bricks_stream_begin": (C.c_int, [P]),
    "r3d_hip_bricks_stream_box": (None, [P, C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_float, C.c_float]),
    "r3d_hip_bricks_stream_point": (None, [P, C.POINTER(C.c_float), C.c_uint32, C.c_float]),
    "r3d_hip_bricks_stream_submit": (C.c_int, [P, P, C.c_uint32, P]),
    "r3d_hip_bricks_stream": (C.c_int, [P, P, C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_float, C.c_float,
                                        C.c_uint32, C.c_uint32, C.c_float, C.c_uint32, P]),
    "r3d_hip_bricks_stream_stats": (C.c_int, [P, C.POINTER(BricksStreamStats)]),
}

_lib = None


def load():
    """Load libr3d_hip.so (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; "
                           f"g.build()'` (make -C render3d_amd/csrc)")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class R3DError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"r3d error {code}: {msg}")
        self.code = code
