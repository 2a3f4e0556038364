"""render3d_amd: MI355X (gfx950) HIP implementation of render3d's two
data-parallel hot paths -- the 3ddct 16^3 chunk decoder and the cube-mode
compute raycaster -- behind the C ABI in include/r3d_hip.h."""
from . import scene
from ._lib import (FrameParams, R3DError, MODE_FULL, MODE_MIP, MODE_DEPTH, MODE_HEATMAP, MODE_RAYDIR,
                   MODE_FLAT, VIEW_ORTHO, VIEW_HALF, R3D_DECODE_EXACT_F32, R3D_DECODE_MFMA_F32, R3D_DECODE_MFMA_F16,
                   R3D_DECODE_MFMA_F16X3, R3D_DECODE_F32_OUT, QUALITY_FULL, QUALITY_FAST, load as load_library,
                   LIB_PATH)
from .api import Bricks, Context, ShardStore, Volume, headers_tensor, headers_from_tensor

__all__ = ["scene", "Bricks", "FrameParams", "R3DError", "Context", "ShardStore", "Volume", "headers_tensor", "headers_from_tensor",
           "load_library", "LIB_PATH", "MODE_FULL", "MODE_MIP", "MODE_DEPTH", "MODE_HEATMAP", "MODE_RAYDIR",
           "MODE_FLAT", "VIEW_ORTHO", "VIEW_HALF", "R3D_DECODE_EXACT_F32", "R3D_DECODE_MFMA_F32", "R3D_DECODE_MFMA_F16",
           "R3D_DECODE_MFMA_F16X3", "R3D_DECODE_F32_OUT", "QUALITY_FULL", "QUALITY_FAST"]
