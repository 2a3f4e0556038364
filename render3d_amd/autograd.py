"""torch.autograd.Function over the differentiable raycast.

forward  -> r3d_hip_raycast_diff  (renders, keeps per-ray checkpoints)
backward -> r3d_hip_raycast_adjoint (hand-written adjoint: dL/dtf, dL/dgrid)

torch only supplies tensors, streams and the optimiser; both passes run in
libr3d_hip.so.
"""
import ctypes as C

import torch

from ._lib import DiffParams, FrameParams, Image
from .api import _stream


class DiffRenderer:
    """Fixed views over one volume; __call__(tf, grid) -> [H, W, 4] image.

    Each autograd call keeps its own checkpoint workspace until its backward
    has run, so several renders of one renderer may enter one loss. The raw
    forward_raw / adjoint_raw pair shares the renderer's workspace: an
    adjoint_raw uses the checkpoints of the latest forward_raw."""

    def __init__(self, ctx, vol, views, width, height, alpha_cutoff=0.98):
        self.ctx = ctx
        self.vol = vol
        self.views = (FrameParams * len(views))(*views)
        self.nviews = len(views)
        self.width = width
        self.height = height
        self.cutoff = alpha_cutoff
        self.workspace_bytes = max(int(ctx.lib.r3d_hip_diff_workspace_bytes(self.views, self.nviews)), 16)
        self.workspace = self.new_workspace()

    def new_workspace(self):
        return torch.empty(self.workspace_bytes, dtype=torch.uint8, device=f"cuda:{self.ctx.device}")

    def params(self, tf, grid, workspace=None, checkpoints=True):
        ws = self.workspace if workspace is None else workspace
        gz, gy, gx = grid.shape
        return DiffParams(tf.data_ptr(), grid.data_ptr(), (C.c_uint32 * 3)(gx, gy, gz), self.cutoff,
                          ws.data_ptr() if checkpoints else None, ws.numel() if checkpoints else 0)

    def forward_raw(self, tf, grid, out=None, workspace=None, checkpoints=True):
        """r3d_hip_raycast_diff; checkpoints=False renders without keeping any
        (no adjoint can follow)."""
        img = out if out is not None else torch.empty((self.height, self.width, 4), dtype=torch.float32,
                                                      device=tf.device)
        im = Image(img.data_ptr(), None, self.width, self.height)
        dp = self.params(tf, grid, workspace, checkpoints)
        self.ctx.check(self.ctx.lib.r3d_hip_raycast_diff(self.ctx.h, self.vol.h, self.views, self.nviews,
                                                         C.byref(im), C.byref(dp), _stream(None)))
        return img

    def adjoint_raw(self, tf, grid, dimg, dtf=None, dgrid=None, workspace=None):
        dimg = dimg.contiguous()
        dtf = dtf if dtf is not None else torch.empty_like(tf)
        if dgrid is None:
            dgrid = torch.zeros_like(grid)
        else:
            dgrid.zero_()
        dp = self.params(tf, grid, workspace)
        self.ctx.check(self.ctx.lib.r3d_hip_raycast_adjoint(
            self.ctx.h, self.vol.h, self.views, self.nviews, C.c_void_p(dimg.data_ptr()), self.width,
            self.height, C.c_void_p(dtf.data_ptr()), C.c_void_p(dgrid.data_ptr()), C.byref(dp), _stream(None)))
        return dtf, dgrid

    def __call__(self, tf, grid):
        return _RaycastFn.apply(tf, grid, self)


class _RaycastFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, tf, grid, renderer):
        tf = tf.detach().contiguous()
        grid = grid.detach().contiguous()
        ws = renderer.new_workspace()  # this call's checkpoints, freed with the autograd node
        img = renderer.forward_raw(tf, grid, workspace=ws)
        ctx.renderer = renderer
        ctx.workspace = ws
        ctx.save_for_backward(tf, grid)
        return img

    @staticmethod
    def backward(ctx, dimg):
        tf, grid = ctx.saved_tensors
        dtf, dgrid = ctx.renderer.adjoint_raw(tf, grid, dimg, workspace=ctx.workspace)
        ctx.workspace = None
        return dtf, dgrid, None
