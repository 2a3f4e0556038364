"""Lane-sharded rays (k_raycast_tiers): a launch of at least 4,096 8x8 tiles
renders its costliest FULL/FAST tiles with a lane pair per ray and the rest
with one lane per ray. Pixels, counters and per-pixel sample counts must
equal the oracle's and the one-lane-per-ray kernel's (another tile shape
takes k_raycast), in every mode the tiers render and with and without
empty-space skips."""
import math

import numpy as np
import pytest

import render3d_amd as r3d
from oracle import oracle as orc
from render3d_amd import scene

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

TOL = 1e-5
W = H = 512  # 4,096 tiles of 8x8: the smallest ordered (tiered) launch


def _views():
    lut = scene.tf_ramp_40_200()
    gate = scene.skip_gate(lut)
    cam = scene.orbit_camera(45.0, 10.0, 2.0)
    base = dict(step_voxels=0.5, gate=gate)
    out = {
        "full": [scene.frame_params(cam, W, H, **base)],
        "flat": [scene.frame_params(cam, W, H, mode=r3d.MODE_FLAT, **base)],
        "mip": [scene.frame_params(cam, W, H, mode=r3d.MODE_MIP, **base)],
        "lodbias": [scene.frame_params(cam, W, H, lod_bias=1.5, frame_index=7, **base)],
        "threshold": [scene.frame_params(cam, W, H, threshold=0.3, **base)],
    }
    ortho = scene.frame_params(cam, W, H, view_flags=r3d.VIEW_ORTHO, **base)
    ortho.cam_right[:] = [x * 0.9 for x in ortho.cam_right]
    out["ortho"] = [ortho]
    rot = scene.frame_params(scene.orbit_camera(-30.0, 25.0, 1.6), W, H, step_voxels=1.0, gate=gate)
    a = math.radians(20)
    rot.vol_r0[:] = [math.cos(a), 0.0, math.sin(a)]
    rot.vol_r1[:] = [0.0, 1.0, 0.0]
    rot.vol_r2[:] = [-math.sin(a), 0.0, math.cos(a)]
    rot.vol_tx, rot.vol_ty, rot.vol_tz = 0.05, -0.02, 0.01
    out["rotated"] = [rot]
    inside = scene.Camera((0.45, 0.55, 0.3))
    inside.yaw = np.float32(0.3)
    out["inside"] = [scene.frame_params(inside, W, H, step_voxels=0.75, gate=gate)]
    return lut, out


def _multi():
    """Four 512^2 views in a 1024^2 image: FULL, MIP and FLAT each with
    4,096 tiles or more in one call."""
    lut = scene.tf_ramp_40_200()
    base = dict(step_voxels=0.5, gate=scene.skip_gate(lut))
    vs = []
    modes = [r3d.MODE_FULL, r3d.MODE_MIP, r3d.MODE_FLAT, r3d.MODE_FULL]
    for i, m in enumerate(modes):
        cam = scene.orbit_camera(30.0 + 80.0 * i, 15.0 * (i % 3 - 1), 1.9 + 0.1 * i)
        vs.append(scene.frame_params(cam, W, H, mode=m, view_org=((i % 2) * W, (i // 2) * H), **base))
    return lut, vs


@pytest.fixture(scope="module")
def dense96(ctx):
    dims = (96, 80, 72)
    vol = orc.synth_sheets(*dims, 4, 31)
    v = ctx.volume(*dims)
    v.upload(vol)
    return v, orc.OracleVolume(vol, *dims)


@pytest.fixture(scope="module")
def sparse96(ctx):
    """Sheets on an empty background: rays skip empty 8^3 blocks."""
    dims = (96, 80, 72)
    vol = orc.synth_sheets(*dims, 2, 31)
    vol[vol < 90] = 0
    v = ctx.volume(*dims)
    v.upload(vol)
    return v, orc.OracleVolume(vol, *dims)


def _render(ctx, v, lut, vs, w, h, quality, tile=(8, 8)):
    ctx.set_transfer(lut)
    ctx.set_quality(quality)
    ctx.set_raycast_tuning(*tile)
    try:
        out = torch.zeros((h, w, 4), dtype=torch.float32, device="cuda")
        res = ctx.raycast(v, vs, w, h, alpha_cutoff=0.99, out=out, px_steps=True, stats=True)
        torch.cuda.synchronize()
    finally:
        ctx.set_quality(r3d.QUALITY_FULL)
        ctx.set_raycast_tuning(8, 8)
    return res["rgba"].cpu().numpy(), res["px_steps"].cpu().numpy().astype(np.uint32), res["stats"]


def _check(ctx, v, ov, lut, vs, w, h, quality, want_skips):
    img, steps, st = _render(ctx, v, lut, vs, w, h, quality)
    ref, cnt, rsteps = orc.raycast(ov, scene.tf_unorm(lut), vs, w, h, alpha_cutoff=0.99, px_steps=True,
                                   quality=orc.QUALITY_FAST if quality == r3d.QUALITY_FAST else orc.QUALITY_FULL)
    err = float(np.max(np.abs(img - ref)))
    print(f"max abs {err:g}, samples {cnt['samples']}, skips {cnt['skips']}, shaded {cnt['shaded']}")
    assert err <= TOL
    assert np.array_equal(steps, rsteps)
    assert (st["samples"], st["skips"], st["shaded"], st["rays"]) == \
        (cnt["samples"], cnt["skips"], cnt["shaded"], cnt["rays"]), (st, cnt)
    if want_skips:
        assert cnt["skips"] > 0
    # one lane per ray (16x8 tiles take k_raycast): the same bytes
    img1, steps1, st1 = _render(ctx, v, lut, vs, w, h, quality, tile=(16, 8))
    assert np.array_equal(img.view(np.uint32), img1.view(np.uint32))
    assert np.array_equal(steps, steps1)
    counts = ("samples", "skips", "shaded", "rays")
    assert [st[k] for k in counts] == [st1[k] for k in counts]


@pytest.mark.parametrize("case", ["full", "fast", "flat", "mip", "lodbias", "threshold", "ortho", "rotated", "inside"])
def test_tiers_match_oracle(ctx, dense96, case):
    v, ov = dense96
    lut, views = _views()
    quality = r3d.QUALITY_FAST if case == "fast" else r3d.QUALITY_FULL
    _check(ctx, v, ov, lut, views["full" if case == "fast" else case], W, H, quality, False)


@pytest.mark.parametrize("case", ["full", "fast", "flat", "mip", "rotated"])
def test_tiers_with_skips_match_oracle(ctx, sparse96, case):
    v, ov = sparse96
    lut, views = _views()
    quality = r3d.QUALITY_FAST if case == "fast" else r3d.QUALITY_FULL
    _check(ctx, v, ov, lut, views["full" if case == "fast" else case], W, H, quality, True)


@pytest.mark.parametrize("case", ["half", "ragged"])
def test_tiers_half_and_ragged_views(ctx, dense96, case):
    """A half-resolution view (each ray fills 2x2 pixels) of 1024^2, and a
    600x520 view whose last tile column and row are partial: both launches
    have 4,096 tiles or more."""
    v, ov = dense96
    lut = scene.tf_ramp_40_200()
    gate = scene.skip_gate(lut)
    cam = scene.orbit_camera(60.0, -15.0, 1.9)
    if case == "half":
        w, h = 2 * W, 2 * H
        vs = [scene.frame_params(cam, w, h, view_flags=r3d.VIEW_HALF, step_voxels=0.5, gate=gate)]
    else:
        w, h = 600, 520
        vs = [scene.frame_params(cam, w, h, step_voxels=0.5, gate=gate)]
    _check(ctx, v, ov, lut, vs, w, h, r3d.QUALITY_FULL, False)


def test_tiers_multi_view(ctx, dense96):
    v, ov = dense96
    lut, vs = _multi()
    _check(ctx, v, ov, lut, vs, 2 * W, 2 * H, r3d.QUALITY_FULL, False)
