"""Closed forms that pin the raycast oracle's shading, mip blend and
empty-space skip (raycast.slang:957-998, vkbackend.c:1058-1069 with
raycast.slang:771-777, raycast.slang:865-868), each with a planted mistake
in the oracle that the same check must reject.

- Shading: on a field linear in x, the central-difference (and the
  tetrahedral) gradient is (2 lod_step b / 255, 0, 0) exactly up to FP32
  rounding, so the normal is +x and the Blinn-Phong terms follow in FP64
  from the ray direction alone. An opaque TF (alpha 1) with a constant colour
  and a low cut make each ray's pixel the shaded colour of its first sample
  above the cut, which lies in the interior.
- Mip blend: hand-made levels of constant value A (level 0) and B (level 1).
  An orthographic ray has lod = lod_bias exactly when cone / voxel = 1, so
  the MIP image is (A + (B - A) lod) / 255; a perspective ray's first sample
  has lod = log2(t0 cone / voxel) and, with A > B, is the ray's maximum.
- Skip jump: one occupied 8^3 block in an empty volume and rays along +z.
  An FP64 march of the reference loop gives every sample's t: after a jump
  the ray stands at the next block face plus 1% of the block. Per-pixel
  iteration counts, the skip count and the first-hit depth (DEPTH mode) must
  agree.
"""
import math

import numpy as np
import pytest

from oracle import oracle as orc
from render3d_amd._lib import FrameParams, MODE_DEPTH, MODE_FULL, MODE_MIP

f32 = np.float32
KEY = np.array([0.4, 0.8, -0.4]) / math.sqrt(0.96)


def manual_view(origin, fwd, right, up, w, h, *, mode=MODE_FULL, ortho=False, step_voxels=0.5, lod_bias=0.0,
                max_mip=10.0, threshold=0.0, gate=0.0):
    """An r3d_frame_params with an explicit camera basis (identity model transform)."""
    p = FrameParams()
    p.cam_origin[:] = [float(v) for v in origin]
    p.cam_forward[:] = [float(v) for v in fwd]
    p.cam_right[:] = [float(v) for v in right]
    p.cam_up[:] = [float(v) for v in up]
    p.step_voxels = step_voxels
    p.density = 1.0
    p.lod_bias = lod_bias
    p.max_mip = max_mip
    p.viewport[0], p.viewport[1] = w, h
    p.mode = mode
    p.threshold = threshold
    p.skip_gate = gate
    p.vol_r0[:] = [1.0, 0.0, 0.0]
    p.vol_r1[:] = [0.0, 1.0, 0.0]
    p.vol_r2[:] = [0.0, 0.0, 1.0]
    p.view_flags = 1 if ortho else 0
    return p


# ---- (a) shading on a linear field --------------------------------------------------

N_RAMP = 32
RAMP_SLOPE = 6  # u8 per voxel along x: 10 .. 196
COLOUR = np.array([0.8, 0.6, 0.4])


def ramp_volume():
    x = np.arange(N_RAMP, dtype=np.int64)
    v = np.broadcast_to(10 + RAMP_SLOPE * x, (N_RAMP, N_RAMP, N_RAMP))
    return np.ascontiguousarray(v, np.uint8).reshape(-1)


def opaque_tf():
    tf = np.zeros((256, 4), np.float32)
    tf[:, :3] = COLOUR.astype(np.float32)
    tf[:, 3] = 1.0
    return tf


def shaded_colour(d, slope, radius):
    """FP64 Blinn-Phong of raycast.slang:985-997 for the normal +x: slope in
    u8 per voxel of the sampled level, tap radius in voxels of that level."""
    d = np.asarray(d, np.float64) / np.linalg.norm(d)
    gm = 2.0 * radius * slope / 255.0  # |g| of two taps 2 * radius apart
    n = np.array([1.0, 0.0, 0.0])
    l = -d
    h = (l - d) / np.linalg.norm(l - d)
    diff = 0.7 * abs(n @ l) + 0.3 * abs(n @ KEY)
    spec = abs(n @ h) ** 24 * 0.3
    lit = 0.25 + 0.75 * diff
    w = min(max(gm * 6.0, 0.0), 1.0)
    return COLOUR * (1.0 + (lit - 1.0) * w) + spec * w


def ramp_render(lod_bias, quality):
    ov = orc.OracleVolume(ramp_volume(), N_RAMP, N_RAMP, N_RAMP)
    fwd = (1.0, 0.15, 0.1)
    # orthographic, 16 x 16 rays entering through the x = 0 face; cone / voxel
    # = 2 |up| / H * 32 = 1, so lod = lod_bias exactly
    p = manual_view((-0.5, 0.45, 0.45), fwd, (0.0, 0.25, 0.0), (0.0, 0.0, 0.25), 16, 16, ortho=True,
                    lod_bias=lod_bias, max_mip=4.0, threshold=100.0 / 255.0)
    img, cnt, _ = orc.raycast(ov, opaque_tf(), [p], 16, 16, alpha_cutoff=0.99, quality=quality)
    return img, cnt, fwd


@pytest.mark.parametrize("quality", [orc.QUALITY_FULL, orc.QUALITY_FAST])
@pytest.mark.parametrize("lod_bias,slope", [(-100.0, RAMP_SLOPE), (1.0, 2 * RAMP_SLOPE)])
def test_shading_closed_form_on_a_linear_field(quality, lod_bias, slope):
    """Level 1 of the ramp is the rounded 2x box 13 + 12 x, again linear. At
    lod 1 the tap radius voxel * lod_step = 2 / 32 is one level-1 voxel, as
    one level-0 voxel at lod 0."""
    img, cnt, fwd = ramp_render(lod_bias, quality)
    want = shaded_colour(fwd, slope, 1.0)
    assert cnt["shaded"] == 256 and cnt["skips"] == 0
    tol = 2e-6 if quality == orc.QUALITY_FULL else 1e-5
    assert np.max(np.abs(img[..., :3] - want)) < tol, (img[0, 0], want)
    assert np.all(img[..., 3] == 1.0)


@pytest.mark.parametrize("plant,lod_bias,quality", [
    (orc.PLANT_SPEC_EXP, -100.0, orc.QUALITY_FULL),
    (orc.PLANT_LIT_LERP, -100.0, orc.QUALITY_FULL),
    (orc.PLANT_KEY_LIGHT, -100.0, orc.QUALITY_FULL),
    (orc.PLANT_TAP_RADIUS, 1.0, orc.QUALITY_FULL),
    (orc.PLANT_FAST_RENORM, -100.0, orc.QUALITY_FAST),
])
def test_shading_check_rejects_planted_mistakes(plant, lod_bias, quality):
    slope = RAMP_SLOPE if lod_bias < 0 else 2 * RAMP_SLOPE
    with orc.planted(plant):
        img, _, fwd = ramp_render(lod_bias, quality)
    assert np.max(np.abs(img[..., :3] - shaded_colour(fwd, slope, 1.0))) > 1e-3


def test_fast_and_full_gradients_agree_on_a_linear_field_but_not_elsewhere():
    """The sqrt(3)/2 renorm gives the tetrahedral gradient the central
    difference's response on a linear field (raycast.slang:920-923); on the
    sheet volume the two qualities render differently."""
    a, _, _ = ramp_render(-100.0, orc.QUALITY_FULL)
    b, _, _ = ramp_render(-100.0, orc.QUALITY_FAST)
    assert np.max(np.abs(a - b)) < 1e-5
    n = 48
    vol = orc.synth_sheets(n, n, n, 4, 3)
    ov = orc.OracleVolume(vol, n, n, n)
    from render3d_amd import scene
    lut = scene.tf_unorm(scene.tf_ramp_40_200())
    p = scene.frame_params(scene.orbit_camera(30.0, 10.0, 1.8), 32, 32, step_voxels=0.5)
    full, cf, _ = orc.raycast(ov, lut, [p], 32, 32, quality=orc.QUALITY_FULL)
    fast, cq, _ = orc.raycast(ov, lut, [p], 32, 32, quality=orc.QUALITY_FAST)
    assert cf["shaded"] > 0 and np.max(np.abs(full - fast)) > 1e-3


# ---- (b) the linear mip blend ----------------------------------------------------------

A_LEV, B_LEV = 200, 60


def two_level_volume():
    n = 32
    lev0 = np.full(n ** 3, A_LEV, np.uint8)
    lev1 = np.full((n // 2) ** 3, B_LEV, np.uint8)
    return orc.OracleVolume(lev0, n, n, n, levels=[((n, n, n), lev0), ((n // 2,) * 3, lev1)])


def mip_ortho(lod_bias):
    p = manual_view((0.5, 0.5, -1.0), (0.0, 0.0, 1.0), (0.25, 0.0, 0.0), (0.0, 0.25, 0.0), 16, 16,
                    mode=MODE_MIP, ortho=True, lod_bias=lod_bias, max_mip=1.0)
    img, _, _ = orc.raycast(two_level_volume(), opaque_tf(), [p], 16, 16)
    return img


@pytest.mark.parametrize("lod", [0.125, 0.3, 0.77, 0.999])
def test_mip_blend_orthographic_lod_is_the_bias(lod):
    img = mip_ortho(lod)
    w = float(f32(lod))
    want = (A_LEV + (B_LEV - A_LEV) * w) / 255.0
    assert np.max(np.abs(img[..., 0] - want)) < 1e-6, (img[0, 0, 0], want)


def test_mip_blend_check_rejects_planted_weight():
    with orc.planted(orc.PLANT_MIP_WEIGHT):
        img = mip_ortho(0.3)
    assert np.max(np.abs(img[..., 0] - (A_LEV + (B_LEV - A_LEV) * float(f32(0.3))) / 255.0)) > 1e-2


def hash12_py(px, py, frame):
    m = 0xFFFFFFFF
    h = (px * 374761393 + py * 668265263 + frame * 2246822519) & m
    h = ((h ^ (h >> 13)) * 1274126177) & m
    return (h >> 8) / 16777216.0


def mip_perspective():
    """Camera at z = -2 looking +z; cone = 2 |up| / H with |up| = 0.15,
    H = 16, so t * cone / voxel = 0.6 t lies in (1.2, 1.9) along every ray."""
    W = H = 16
    p = manual_view((0.5, 0.5, -2.0), (0.0, 0.0, 1.0), (0.15, 0.0, 0.0), (0.0, 0.15, 0.0), W, H, mode=MODE_MIP,
                    max_mip=1.0)
    img, _, _ = orc.raycast(two_level_volume(), opaque_tf(), [p], W, H)
    want = np.zeros((H, W))
    for py in range(H):
        for px in range(W):
            ndx, ndy = (px + 0.5) / W * 2 - 1, 1 - (py + 0.5) / H * 2
            d = np.array([0.15 * ndx, 0.15 * ndy, 1.0])
            d /= np.linalg.norm(d)
            tn = 2.0 / d[2]  # enters through the z = 0 face
            t0 = tn + hash12_py(px, py, 0) * 0.5 / 32
            lod = math.log2(t0 * (2 * 0.15 / H) * 32)
            assert 0 < lod < 1
            want[py, px] = (A_LEV + (B_LEV - A_LEV) * lod) / 255.0
    return img[..., 0], want


def test_mip_blend_perspective_lod_from_the_ray_cone():
    got, want = mip_perspective()
    assert np.max(np.abs(got - want)) < 2e-6


def test_mip_blend_perspective_rejects_planted_weight():
    with orc.planted(orc.PLANT_MIP_WEIGHT):
        got, want = mip_perspective()
    assert np.max(np.abs(got - want)) > 1e-2


# ---- (c) the occupancy skip jump --------------------------------------------------------

N_SKIP = 64
BLOCK_LO, BLOCK_HI = 32, 40  # the one occupied 8^3 block (voxels)


def skip_volume():
    v = np.zeros((N_SKIP,) * 3, np.uint8)
    v[BLOCK_LO:BLOCK_HI, BLOCK_LO:BLOCK_HI, BLOCK_LO:BLOCK_HI] = 200
    return v.reshape(-1)


SKIP_W = SKIP_H = 4
SKIP_GATE = 10.0 / 255.0
SKIP_ORG = (0.56, 0.56, -0.5)
SKIP_RIGHT = 0.05  # rays at x, y in [0.5225, 0.5975]: inside the block's column


def skip_render():
    ov = orc.OracleVolume(skip_volume(), N_SKIP, N_SKIP, N_SKIP)
    p = manual_view(SKIP_ORG, (0.0, 0.0, 1.0), (SKIP_RIGHT, 0.0, 0.0), (0.0, SKIP_RIGHT, 0.0), SKIP_W, SKIP_H,
                    mode=MODE_DEPTH, ortho=True, lod_bias=-100.0, gate=SKIP_GATE)
    img, cnt, steps = orc.raycast(ov, opaque_tf(), [p], SKIP_W, SKIP_H, px_steps=True)
    return img, cnt, steps


def skip_march_fp64(px, py):
    """The reference loop (raycast.slang:760-869, 1020) in FP64 for one ray
    along +z through the block's column: returns (iterations, skips,
    first-hit t, smallest distance of a decision to its threshold)."""
    n, bs, dt = N_SKIP, 8.0 / N_SKIP, 0.5 / N_SKIP
    ndx, ndy = (px + 0.5) / SKIP_W * 2 - 1, 1 - (py + 0.5) / SKIP_H * 2
    x, y, z0 = SKIP_ORG[0] + SKIP_RIGHT * ndx, SKIP_ORG[1] + SKIP_RIGHT * ndy, SKIP_ORG[2]
    inside_xy = all(BLOCK_LO <= int(c * n - 0.5) and int(c * n - 0.5) + 1 < BLOCK_HI for c in (x, y))
    assert inside_xy
    gate = max(float(f32(SKIP_GATE)), 0.5 / 255)
    tn, tfar = -z0, 1.0 - z0
    t = tn + hash12_py(px, py, 0) * dt
    steps = skips = 0
    first_hit = -1.0
    margin = 1.0
    while t < tfar and steps < 4096:
        margin = min(margin, tfar - t)
        steps += 1
        pz = z0 + t
        tz = min(max(pz * n - 0.5, 0.0), n - 1.0)
        i0 = int(tz)
        f = tz - i0
        occ_in = lambda i: BLOCK_LO <= i < BLOCK_HI
        v = 200.0 * ((1 - f) * occ_in(i0) + f * occ_in(min(i0 + 1, n - 1))) / 255.0
        blk = min(max(int(math.floor(pz * 8)), 0), 7)
        occ = 200.0 / 255.0 if 3 <= blk <= 5 else 0.0  # the block max, dilated by one block
        if v > 0:
            margin = min(margin, abs(v - gate), abs(v - 0.05))
        if 0 < round(pz * 8) < 8:  # an interior block face (the grid clamps at 0 and 8)
            margin = min(margin, abs(pz * 8 - round(pz * 8)))
        if v < gate and occ < gate:
            nextz = (math.floor(pz / bs) + 1.0) * bs
            t = max(nextz - z0, t) + bs * 0.01
            skips += 1
            continue
        if first_hit < 0 and v > 0.05:
            first_hit = t
        t += dt
    return steps, skips, first_hit, margin


def test_skip_jump_lands_one_percent_past_the_next_face():
    img, cnt, steps = skip_render()
    tot_skips, checked = 0, 0
    for py in range(SKIP_H):
        for px in range(SKIP_W):
            ns, nk, fh, margin = skip_march_fp64(px, py)
            tot_skips += nk
            assert margin > 1e-5, "a decision lies within rounding of its threshold: pick another frame"
            assert nk == 5  # blocks 0, 1, 2 before the occupied column and 6, 7 after it
            assert steps[py, px] == ns, (px, py, steps[py, px], ns)
            assert abs(img[py, px, 0] - (1.0 - fh / 2.0)) < 5e-6, (px, py, img[py, px, 0], 1 - fh / 2)
            checked += 1
    assert checked == SKIP_W * SKIP_H
    assert cnt["skips"] == tot_skips and cnt["samples"] == int(steps.sum())


@pytest.mark.parametrize("plant", [orc.PLANT_SKIP_NUDGE, orc.PLANT_SKIP_FACE])
def test_skip_check_rejects_planted_mistakes(plant):
    with orc.planted(plant):
        img, cnt, steps = skip_render()
    bad = 0
    for py in range(SKIP_H):
        for px in range(SKIP_W):
            ns, nk, fh, _ = skip_march_fp64(px, py)
            bad += steps[py, px] != ns or abs(img[py, px, 0] - (1.0 - fh / 2.0)) > 5e-6
    assert bad > 0
