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
    """Fixed views over one volume; __call__(tf, grid) -> [H, W, 4] image."""

    def __init__(self, ctx, vol, views, width, height, alpha_cutoff=0.98):
        self.ctx = ctx
        self.vol = vol
        self.views = (FrameParams * len(views))(*views)
        self.nviews = len(views)
        self.width = width
        self.height = height
        self.cutoff = alpha_cutoff
        nbytes = ctx.lib.r3d_hip_diff_workspace_bytes(self.views, self.nviews)
        self.workspace = torch.empty(max(int(nbytes), 16), dtype=torch.uint8, device=f"cuda:{ctx.device}")

    def params(self, tf, grid):
        gz, gy, gx = grid.shape
        return DiffParams(tf.data_ptr(), grid.data_ptr(), (C.c_uint32 * 3)(gx, gy, gz), self.cutoff,
                          self.workspace.data_ptr(), self.workspace.numel())

    def forward_raw(self, tf, grid, out=None):
        img = out if out is not None else torch.empty((self.height, self.width, 4), dtype=torch.float32,
                                                      device=tf.device)
        im = Image(img.data_ptr(), None, self.width, self.height)
        dp = self.params(tf, grid)
        self.ctx.check(self.ctx.lib.r3d_hip_raycast_diff(self.ctx.h, self.vol.h, self.views, self.nviews,
                                                         C.byref(im), C.byref(dp), _stream(None)))
        return img

    def adjoint_raw(self, tf, grid, dimg, dtf=None, dgrid=None):
        dimg = dimg.contiguous()
        dtf = dtf if dtf is not None else torch.empty_like(tf)
        if dgrid is None:
            dgrid = torch.zeros_like(grid)
        else:
            dgrid.zero_()
        dp = self.params(tf, grid)
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
        img = renderer.forward_raw(tf, grid)
        ctx.renderer = renderer
        ctx.save_for_backward(tf, grid)
        return img

    @staticmethod
    def backward(ctx, dimg):
        tf, grid = ctx.saved_tensors
        dtf, dgrid = ctx.renderer.adjoint_raw(tf, grid, dimg)
        return dtf, dgrid, None
