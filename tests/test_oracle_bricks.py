"""The bricks-LOD oracle (oracle_raycast.c: r3do_raycast_bricks,
r3do_bricks_build) pinned without a runnable reference shader:

* page table and atlas against a numpy restatement of the layout
  (vkbackend.c:4185-4205) over the pinned pyramid;
* inside one brick the bricks fetch is the cube path's trilinear, so MIP,
  DEPTH, FLAT and HEATMAP renders of a one-brick level equal the cube oracle's
  bit for bit (level 0 of 128^3, and level 1 of 256^3 with level 0 absent);
  FULL differs only in the rounding of its gradient taps;
* closed forms on hand-made page tables of constant slots: the level the
  locate resolves for every desired level and residency
  (raycast.slang:198-247), the per-ray cache that keeps the brick it found
  while the ray stays inside it (raycast.slang:255-279), whole-brick MIP skips,
  zoom crop and slabs (raycast.slang:690-718) and a non-cubic shape.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as orc
from render3d_amd import scene
from render3d_amd._lib import (BR_LOD_FLAG, FrameParams, MODE_DEPTH, MODE_FLAT, MODE_FULL, MODE_HEATMAP,
                               MODE_MIP)

f32 = np.float32
K255 = f32(1.0) / f32(255.0)
BG = np.array([0.02, 0.02, 0.03, 1.0], f32)
ABSENT = 0xFFFFFFFF


def level_dims(shape, nlev):
    d = [tuple(shape)]
    for _ in range(1, nlev):
        d.append(tuple(max(a // 2, 1) for a in d[-1]))
    return d


class Page:
    """A hand-made page table (header as vkbackend.c:4185-4205, every entry absent)."""

    def __init__(self, shape, nlev, abpa):
        self.dims = level_dims(shape, nlev)
        self.nb = [tuple(-(-a // 128) for a in d) for d in self.dims]
        self.words = orc.brick_page_words(self.dims)
        self.page = np.full(self.words, ABSENT, np.uint32)
        self.page[:64] = 0
        self.page[0] = nlev
        self.page[1:4] = shape
        self.off = []
        w = 64
        for l, (bx, by, bz) in enumerate(self.nb):
            self.page[4 + 4 * l:7 + 4 * l] = (w, bx | by << 10 | bz << 20, 1 << l)
            self.off.append(w)
            w += bx * by * bz
        self.abpa = abpa
        self.atlas = np.zeros(abpa ** 3 * orc.SLOT_VOX, np.uint8)
        self.slots = {}

    def slot(self, value):
        """A slot of constant `value` (shared by every entry that uses it)."""
        if value not in self.slots:
            s = len(self.slots)
            self.atlas[s * orc.SLOT_VOX:(s + 1) * orc.SLOT_VOX] = value
            self.slots[value] = s
        return self.slots[value]

    def put(self, l, bx, by, bz, value):
        nbx, nby, _ = self.nb[l]
        self.page[self.off[l] + (bz * nby + by) * nbx + bx] = self.slot(value) | value << 24

    def fill(self, l, value):
        for bz in range(self.nb[l][2]):
            for by in range(self.nb[l][1]):
                for bx in range(self.nb[l][0]):
                    self.put(l, bx, by, bz, value)

    def bricks(self):
        return orc.OracleBricks(self.page, self.atlas, self.abpa)


def z_view(W, H, abpa, mode=MODE_MIP, bias=0.0, flags=0, threshold=0.0):
    """Orthographic view straight down +z over the unit square: pixel (px, py)
    casts the ray x = (px + 0.5) / W, y = 1 - (py + 0.5) / H."""
    p = FrameParams()
    p.cam_origin[:] = [0.5, 0.5, -1.0]
    p.cam_right[:] = [0.5, 0.0, 0.0]
    p.cam_up[:] = [0.0, 0.5, 0.0]
    p.cam_forward[:] = [0.0, 0.0, 1.0]
    p.step_voxels = 1.0
    p.density = 1.0
    p.lod_bias = bias
    p.max_mip = 10.0
    p.viewport[0], p.viewport[1] = W, H
    p.mode = mode
    p.threshold = threshold
    p.vol_r0[:] = [1.0, 0.0, 0.0]
    p.vol_r1[:] = [0.0, 1.0, 0.0]
    p.vol_r2[:] = [0.0, 0.0, 1.0]
    p.view_flags = 1 | flags
    p.brick_mode = abpa | BR_LOD_FLAG
    return p


def render(ob, views, W, H, lut=None, **kw):
    lut = scene.tf_unorm(scene.tf_ramp_40_200()) if lut is None else lut
    return orc.raycast_bricks(ob, lut, views if isinstance(views, list) else [views], W, H, **kw)


def mip_of(value):
    return np.float32(f32(value) * K255)


# ---- page table and atlas ----------------------------------------------------

def test_build_matches_numpy_layout():
    """Non-cubic shape, dims not multiples of 128, three levels, some bricks
    not resident: header, entry order, slot order, brick maxima and the zero
    padding of edge bricks against numpy."""
    nx, ny, nz = 300, 200, 150
    vol = orc.synth_sheets(nx, ny, nz, 12, 3)
    ov = orc.OracleVolume(vol, nx, ny, nz, nlev=3)
    dims = [d for d, _ in ov.levels]
    nent = orc.brick_page_words(dims) - 64
    resident = np.ones(nent, np.uint8)
    resident[[1, 4, 12]] = 0
    ob = orc.bricks_build(ov, 3, resident)
    pg = ob.page
    assert pg[0] == 3 and tuple(pg[1:4]) == (nx, ny, nz)
    w, slot = 64, 0
    for l, ((lx, ly, lz), arr) in enumerate(ov.levels):
        nb = (-(-lx // 128), -(-ly // 128), -(-lz // 128))
        assert tuple(pg[4 + 4 * l:7 + 4 * l]) == (w, nb[0] | nb[1] << 10 | nb[2] << 20, 1 << l)
        pad = np.zeros((nb[2] * 128, nb[1] * 128, nb[0] * 128), np.uint8)
        pad[:lz, :ly, :lx] = arr.reshape(lz, ly, lx)
        for bz in range(nb[2]):
            for by in range(nb[1]):
                for bx in range(nb[0]):
                    e = pg[w]
                    if not resident[w - 64]:
                        assert e == ABSENT
                    else:
                        blk = pad[bz * 128:bz * 128 + 128, by * 128:by * 128 + 128, bx * 128:bx * 128 + 128]
                        assert e == slot | int(blk.max()) << 24
                        got = ob.atlas[slot * orc.SLOT_VOX:(slot + 1) * orc.SLOT_VOX].reshape(128, 128, 128)
                        assert np.array_equal(got, blk), (l, bx, by, bz)
                        slot += 1
                    w += 1
    assert w == pg.size and slot == nent - 3
    assert not ob.atlas[slot * orc.SLOT_VOX:].any()
    with pytest.raises(ValueError):
        orc.bricks_build(ov, 2)  # 15 bricks, 8 slots


# ---- one brick: the cube path's sampling -------------------------------------

def ripple(n):
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    f = 128.0 + 70.0 * np.sin(x * 0.13 + 0.5) * np.cos(y * 0.07) + 30.0 * np.sin(z * 0.11 + y * 0.05)
    return np.clip(f + 0.5, 12, 250).astype(np.uint8).reshape(-1)


def views_for(n_views=3, **kw):
    """Cube-mode views: perspective orbit, orthographic, and a rotated model transform."""
    a = scene.frame_params(scene.orbit_camera(30.0, 20.0, 1.9), 48, 40, **kw)
    b = scene.frame_params(scene.orbit_camera(-60.0, -10.0, 1.8), 48, 40, view_flags=1, **kw)
    c = scene.frame_params(scene.orbit_camera(100.0, 35.0, 2.2), 48, 40, frame_index=7, **kw)
    ca, sa = np.cos(0.4), np.sin(0.4)
    c.vol_r0[:] = [ca, 0.0, sa]
    c.vol_r1[:] = [0.0, 1.0, 0.0]
    c.vol_r2[:] = [-sa, 0.0, ca]
    c.vol_tx, c.vol_ty = 0.05, -0.03
    return [a, b, c][:n_views]


def as_bricks(v, abpa):
    b = FrameParams.from_buffer_copy(v)
    b.brick_mode = abpa | BR_LOD_FLAG
    return b


@pytest.mark.parametrize("mode", [MODE_MIP, MODE_DEPTH, MODE_FLAT, MODE_HEATMAP])
def test_one_brick_level_zero_equals_cube_path(mode):
    n = 128
    ov = orc.OracleVolume(ripple(n), n, n, n, nlev=1)
    ob = orc.bricks_build(ov, 1)
    lut = scene.tf_unorm(scene.tf_ramp_40_200())
    shorter = 0
    for v in views_for(mode=mode, step_voxels=0.8):
        want, _, ws = orc.raycast(ov, lut, [v], 48, 40, px_steps=True)
        got, cnt, gs = orc.raycast_bricks(ob, lut, [as_bricks(v, 1)], 48, 40, px_steps=True)
        assert np.array_equal(got, want)
        if mode == MODE_MIP:  # a sample at the brick's maximum ends the ray, the cube path marches on
            assert (gs <= ws).all()
            shorter += int((gs < ws).sum())
        else:
            assert np.array_equal(gs, ws)
        assert (got[..., 0] != BG[0]).sum() > 300
    assert shorter > 0 or mode != MODE_MIP


@pytest.mark.parametrize("mode", [MODE_MIP, MODE_FLAT, MODE_HEATMAP])
def test_one_brick_level_one_equals_cube_path(mode):
    """256^3 with level 0 absent: lod_bias 1 makes every sample ask for level
    1 (one brick of scale 2), which the cube path samples the same way."""
    n = 256
    ov = orc.OracleVolume(ripple(n), n, n, n, nlev=2)
    resident = np.zeros(9, np.uint8)
    resident[8] = 1
    ob = orc.bricks_build(ov, 1, resident)
    assert (ob.page[64:72] == ABSENT).all()
    lut = scene.tf_unorm(scene.tf_ramp_40_200())
    for v in views_for(n_views=2, mode=mode, lod_bias=1.0, max_mip=1.0):
        want, _, _ = orc.raycast(ov, lut, [v], 48, 40)
        got, _, _ = orc.raycast_bricks(ob, lut, [as_bricks(v, 1)], 48, 40)
        assert np.array_equal(got, want)


def test_one_brick_full_mode_close_to_cube_path():
    """Gradient taps: the cube path offsets the normalised position by
    voxel * lod_step, the bricks path the brick position by 2^lod / scale; the
    two round differently, nothing else differs."""
    n = 128
    ov = orc.OracleVolume(ripple(n), n, n, n, nlev=1)
    ob = orc.bricks_build(ov, 1)
    lut = scene.tf_unorm(scene.tf_ramp_40_200())
    for v in views_for(mode=MODE_FULL):
        want, wc, _ = orc.raycast(ov, lut, [v], 48, 40)
        got, gc, _ = orc.raycast_bricks(ob, lut, [as_bricks(v, 1)], 48, 40)
        assert gc["shaded"] == wc["shaded"] > 1000
        assert np.max(np.abs(got - want)) < 2e-3
        assert np.mean(np.abs(got - want)) < 1e-5


# ---- locate order, cache, skips: constant slots --------------------------------

VAL = (30, 90, 150, 210)  # level l's slots hold VAL[l]


def desired_bias(k, lg):
    """lod_bias that makes lod = k + 0.5 (or the top level) when log2(footprint / voxel) = lg."""
    return k - lg + 0.5


@pytest.mark.parametrize("k", [0, 1, 2, 3])
def test_locate_tries_desired_then_two_coarser_then_the_coarsest(k):
    """512^3, four levels. 16x16 z view: footprint / voxel = 32, log2 5."""
    for present in [(0, 1, 2, 3), (1, 2, 3), (2, 3), (3,), (0,), (1,), (2,), (0, 1, 2), (0, 2), (0, 3)]:
        pg = Page((512, 512, 512), 4, 2)
        for l in present:
            pg.fill(l, VAL[l])
        cand = [l for l in (k, k + 1, k + 2) if l <= 3] + ([3] if k + 2 < 3 else [])
        hit = next((l for l in cand if l in present), None)
        img, _, _ = render(pg.bricks(), z_view(16, 16, 2, bias=desired_bias(k, 5)), 16, 16)
        want = mip_of(VAL[hit]) if hit is not None else f32(0.0)
        assert (img[..., :3] == want).all(), (k, present, hit)


def test_mip_is_max_over_the_column_of_bricks():
    """256^3, level 0 fully resident with a distinct constant per brick:
    the MIP of the z ray through column (bx, by) is its largest constant."""
    pg = Page((256, 256, 256), 2, 3)
    vals = np.array([[[17, 200], [60, 45]], [[120, 33], [90, 240]]])  # [bz][by][bx]
    for bz in range(2):
        for by in range(2):
            for bx in range(2):
                pg.put(0, bx, by, bz, int(vals[bz, by, bx]))
    pg.fill(1, 7)
    img, cnt, _ = render(pg.bricks(), z_view(16, 16, 3, bias=desired_bias(0, 4)), 16, 16)
    for py in range(16):
        for px in range(16):
            bx, by = px // 8, int((256 - (py + 0.5) * 16) // 128)
            assert img[py, px, 0] == mip_of(vals[:, by, bx].max()), (px, py)
    assert cnt["skips"] > 0


def test_ray_cache_keeps_the_coarser_brick():
    """256^3, desired level 0. Column (0, 0): level-0 brick z0 absent, z1 = 200;
    column (1, 0): z0 = 20, z1 = 200; level 1 = 50 everywhere. With threshold
    100/255 samples of 20 and 50 are cut and nothing is skipped. Column (0, 0)
    found level 1 in its first half and the cache keeps that brick for the
    whole ray (its level-1 brick coordinates do not change): 0. Column (1, 0)
    cached a level-0 brick, whose coordinates change at z = 128: 200."""
    pg = Page((256, 256, 256), 2, 2)
    pg.put(0, 0, 0, 1, 200)
    pg.put(0, 1, 0, 0, 20)
    pg.put(0, 1, 0, 1, 200)
    pg.fill(1, 50)
    for thr, want00 in ((f32(100.0) / f32(255.0), f32(0.0)), (f32(0.0), mip_of(50))):
        img, _, _ = render(pg.bricks(), z_view(16, 16, 2, bias=desired_bias(0, 4), threshold=float(thr)), 16, 16)
        col00 = img[8:, :8, 0]  # y < 128, x < 128
        col10 = img[8:, 8:, 0]
        assert (col00 == want00).all(), thr
        assert (col10 == mip_of(200)).all(), thr
        assert (img[:8, :, 0] == (f32(0.0) if thr > 0 else mip_of(50))).all()


def z_bricks():
    """512^3, one level, level-0 brick bz holds 40 * (bz + 1)."""
    pg = Page((512, 512, 512), 1, 2)
    for bz in range(4):
        for by in range(4):
            for bx in range(4):
                pg.put(0, bx, by, bz, 40 * (bz + 1))
    return pg.bricks()


def test_axis_and_oblique_slabs():
    ob = z_bricks()
    for z0, depth, want in ((130, 100, 80), (200, 100, 120), (10, 600, 160), (400, 50, 160)):
        v = z_view(16, 16, 2)
        v.slab_z0, v.slab_depth = z0, depth  # z axis (view_flags >> 8 == 0)
        img, _, _ = render(ob, v, 16, 16)
        assert (img[..., 0] == mip_of(want)).all(), (z0, depth)
        o = z_view(16, 16, 2, flags=4)  # oblique: t range, the volume starts at t = 1
        o.slab_z0, o.slab_depth = 512 + z0, depth
        img, _, _ = render(ob, o, 16, 16)
        assert (img[..., 0] == mip_of(want)).all(), ("oblique", z0, depth)
    v = z_view(16, 16, 2, flags=1 << 8)  # x axis slab: x in [130, 230] voxels
    v.slab_z0, v.slab_depth = 130, 100
    img, _, _ = render(ob, v, 16, 16)
    hit = np.zeros(16, bool)
    hit[4:7] = True  # x = (px + 0.5) * 32: 144, 176, 208
    assert (img[:, hit, 0] == mip_of(160)).all()
    assert (img[:, ~hit] == BG).all()


def test_zoom_crop_box():
    ob = z_bricks()
    v = z_view(16, 16, 2, flags=16)
    v.slab_x0, v.slab_y0, v.slab_z0 = 140.0, 0.0, 300.0
    v.slab_px, v.slab_py, v.slab_nx = 95.0, 300.0, 60.0  # x 140..235, y 0..300, z 300..360
    img, _, _ = render(ob, v, 16, 16)
    ys = 512 - (np.arange(16) + 0.5) * 32
    xs = (np.arange(16) + 0.5) * 32
    hit = (ys[:, None] < 300) & (xs[None, :] >= 140) & (xs[None, :] <= 235)
    assert hit.sum() == 27
    assert (img[hit, 0] == mip_of(120)).all()
    assert (img[~hit] == BG).all()


def test_non_cubic_extent():
    """x 384, y 256, z 128: the volume spans [0,1] x [0,2/3] x [0,1/3]."""
    pg = Page((384, 256, 128), 1, 2)
    for by in range(2):
        for bx in range(3):
            pg.put(0, bx, by, 0, 20 + 30 * (by * 3 + bx))
    img, _, _ = render(pg.bricks(), z_view(24, 24, 2), 24, 24)
    for py in range(24):
        yv = 384 - (py + 0.5) * 16
        for px in range(24):
            if yv >= 256:
                assert (img[py, px] == BG).all()
            else:
                bx, by = int((px + 0.5) * 16 // 128), int(yv // 128)
                assert img[py, px, 0] == mip_of(20 + 30 * (by * 3 + bx)), (px, py)


def test_unsupported_views_raise():
    pg = Page((128, 128, 128), 1, 1)
    pg.fill(0, 9)
    ob = pg.bricks()
    for bad in (dict(brick_mode=2 | BR_LOD_FLAG), dict(brick_mode=1), dict(slab_grid=1), dict(clip_valid=1)):
        v = z_view(8, 8, 1)
        for k, val in bad.items():
            setattr(v, k, val)
        with pytest.raises(ValueError):
            render(ob, v, 8, 8)
