"""Bricks-LOD volumes on the GPU through the C ABI: the page table and slots
r3d_hip_bricks_build writes against the oracle's r3do_bricks_build (exact),
and r3d_hip_bricks_raycast against r3do_raycast_bricks (max-abs <= 1e-5 in
FP32, identical per-pixel sample counts and counters) over every mode, crop
and slabs, orthographic and rotated views, a non-cubic partly resident
pyramid and page tables written by hand as a streamer would."""
import math

import numpy as np
import pytest

import render3d_amd as r3d
from oracle import oracle as orc
from render3d_amd import scene

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

TOL = 1e-5
DIMS = (320, 288, 256)  # x, y, z: 18 + 4 + 1 bricks over three levels
NLEV, ABPA = 3, 3
ABSENT_ENTRIES = [0, 5, 7, 11, 19]


@pytest.fixture(scope="module")
def bscene(ctx):
    nx, ny, nz = DIMS
    vol = orc.synth_sheets(nx, ny, nz, 6, 41)
    ov = orc.OracleVolume(vol, nx, ny, nz, nlev=NLEV)
    nent = orc.brick_page_words([d for d, _ in ov.levels]) - 64
    resident = np.ones(nent, np.uint8)
    resident[ABSENT_ENTRIES] = 0
    ob = orc.bricks_build(ov, ABPA, resident)
    b = r3d.Bricks(ctx, nx, ny, nz, NLEV, ABPA)
    b.build(torch.from_numpy(vol).cuda(), resident)
    torch.cuda.synchronize()
    return b, ob, int(resident.sum())


def test_build_matches_oracle(bscene):
    b, ob, nres = bscene
    assert np.array_equal(b.read_page(), ob.page)
    for s in range(ABPA ** 3):
        want = ob.atlas[s * orc.SLOT_VOX:(s + 1) * orc.SLOT_VOX].reshape(128, 128, 128)
        assert np.array_equal(b.read_slot(s), want), s
    assert nres == 23 - len(ABSENT_ENTRIES)


def _views(W, H, bm):
    lut = scene.tf_ramp_40_200()
    gate = scene.skip_gate(lut)
    cam = scene.orbit_camera(40.0, 15.0, 2.0)
    base = dict(step_voxels=0.5, gate=gate, brick_mode=bm)
    fp = scene.frame_params
    out = {
        "full": [fp(cam, W, H, **base)],
        "flat": [fp(cam, W, H, mode=r3d.MODE_FLAT, **base)],
        "mip": [fp(cam, W, H, mode=r3d.MODE_MIP, **base)],
        "depth": [fp(cam, W, H, mode=r3d.MODE_DEPTH, **base)],
        "heatmap": [fp(cam, W, H, mode=r3d.MODE_HEATMAP, **base)],
        "raydir": [fp(cam, W, H, mode=r3d.MODE_RAYDIR, **base)],
        "lodbias": [fp(cam, W, H, lod_bias=1.3, frame_index=5, **base)],
        "lodbias_neg": [fp(cam, W, H, lod_bias=-2.0, mode=r3d.MODE_MIP, **base)],
        "threshold": [fp(cam, W, H, threshold=0.3, **base)],
        "half": [fp(cam, W, H, view_flags=r3d.VIEW_HALF, **base)],
        "fine_step": [fp(scene.orbit_camera(75.0, -25.0, 1.7), W, H, step_voxels=0.25, gate=gate, brick_mode=bm)],
    }
    ortho = fp(cam, W, H, view_flags=r3d.VIEW_ORTHO, **base)
    ortho.cam_right[:] = [x * 0.9 for x in ortho.cam_right]
    out["ortho"] = [ortho]
    rot = fp(scene.orbit_camera(-30.0, 25.0, 1.6), W, H, step_voxels=1.0, gate=gate, brick_mode=bm)
    a = math.radians(20)
    rot.vol_r0[:] = [math.cos(a), 0.0, math.sin(a)]
    rot.vol_r1[:] = [0.0, 1.0, 0.0]
    rot.vol_r2[:] = [-math.sin(a), 0.0, math.cos(a)]
    rot.vol_tx, rot.vol_ty, rot.vol_tz = 0.05, -0.02, 0.01
    out["rotated"] = [rot]
    inside = scene.Camera((0.45, 0.4, 0.3))
    inside.yaw = np.float32(0.3)
    out["inside"] = [fp(inside, W, H, step_voxels=0.75, gate=gate, brick_mode=bm)]
    crop = fp(cam, W, H, view_flags=16, **base)
    crop.slab_x0, crop.slab_y0, crop.slab_z0 = 40.0, 30.0, 20.0
    crop.slab_px, crop.slab_py, crop.slab_nx = 200.0, 150.0, 180.0
    out["crop"] = [crop]
    for name, flags, z0, depth in (("slab_z", 0, 60.0, 90), ("slab_x", 1 << 8, 100.0, 64), ("slab_y", 2 << 8, 10.0, 200),
                                   ("oblique", 4, 420.0, 150)):
        v = fp(cam, W, H, view_flags=flags, mode=r3d.MODE_MIP if name == "slab_x" else r3d.MODE_FULL, **base)
        v.slab_z0, v.slab_depth = z0, depth
        out[name] = [v]
    w2, h2 = W // 2, H // 2
    out["multi"] = [
        fp(cam, w2, h2, view_org=(0, 0), **base),
        fp(scene.orbit_camera(120.0, -20.0, 2.2), w2, h2, mode=r3d.MODE_MIP, view_org=(w2, 0), **base),
        fp(scene.orbit_camera(200.0, 5.0, 1.8), w2, h2, mode=r3d.MODE_FLAT, view_org=(0, h2), **base),
        fp(scene.orbit_camera(300.0, 40.0, 2.0), w2 - 3, h2 - 5, view_org=(w2, h2), view_flags=16, **base),
    ]
    out["multi"][3].slab_px, out["multi"][3].slab_py, out["multi"][3].slab_nx = 320.0, 100.0, 256.0
    return lut, out


CASES = ["full", "flat", "mip", "depth", "heatmap", "raydir", "lodbias", "lodbias_neg", "threshold", "half",
         "fine_step", "ortho", "rotated", "inside", "crop", "slab_z", "slab_x", "slab_y", "oblique", "multi"]


def check(ctx, b, ob, lut, vs, W, H, quality=r3d.QUALITY_FULL):
    ctx.set_transfer(lut)
    out = torch.zeros((H, W, 4), dtype=torch.float32, device="cuda")
    ctx.set_quality(quality)
    try:
        res = b.raycast(vs, W, H, alpha_cutoff=0.99, out=out, px_steps=True, stats=True, rgba8=True)
    finally:
        ctx.set_quality(r3d.QUALITY_FULL)
    img = res["rgba"].cpu().numpy()
    ref, cnt, steps = orc.raycast_bricks(ob, scene.tf_unorm(lut), vs, W, H, alpha_cutoff=0.99, px_steps=True,
                                         quality=quality)
    err = float(np.max(np.abs(img - ref)))
    assert err <= TOL, f"max abs {err}"
    assert np.array_equal(res["px_steps"].cpu().numpy().astype(np.uint32), steps)
    st = res["stats"]
    assert (st["samples"], st["skips"], st["shaded"], st["rays"]) == \
        (cnt["samples"], cnt["skips"], cnt["shaded"], cnt["rays"]), (st, cnt)
    want8 = (np.clip(ref, 0, 1) * np.float32(255) + np.float32(0.5)).astype(np.uint8)
    want8[ref[..., 3] == 0] = 0
    assert np.array_equal(res["rgba8"].cpu().numpy(), want8)
    return ref, cnt


@pytest.mark.parametrize("case", ["full", "lodbias", "half", "rotated", "crop", "oblique", "multi"])
def test_bricks_fast_quality_matches_oracle(ctx, bscene, case):
    """R3D_QUALITY_FAST: the four tetrahedral taps inside the located brick
    (raycast.slang:933-937), against the oracle's restatement."""
    b, ob, _ = bscene
    W, H = 120, 88
    lut, views = _views(W, H, b.brick_mode)
    _, cnt = check(ctx, b, ob, lut, views[case], W, H, quality=r3d.QUALITY_FAST)
    assert cnt["shaded"] > 0


@pytest.mark.parametrize("case", CASES)
def test_bricks_raycast_matches_oracle(ctx, bscene, case):
    b, ob, _ = bscene
    W, H = 120, 88
    lut, views = _views(W, H, b.brick_mode)
    ref, cnt = check(ctx, b, ob, lut, views[case], W, H)
    print(f"{case}: samples {cnt['samples']} skips {cnt['skips']}")
    if case not in ("raydir",):
        assert cnt["samples"] > 1000


def test_every_tile_shape(ctx, bscene):
    b, ob, _ = bscene
    W, H = 96, 72
    lut, views = _views(W, H, b.brick_mode)
    try:
        for tw, th in ((8, 8), (16, 4), (16, 8), (16, 16)):
            ctx.set_raycast_tuning(tw, th)
            check(ctx, b, ob, lut, views["full"] + [], W, H)
    finally:
        ctx.set_raycast_tuning(8, 8)


def test_streamer_written_tables(ctx):
    """A page table and constant slots written from the host (the way a
    streamer fills them): the per-ray cache case of 256^3 with two levels, and
    an entry whose slot lies past the atlas, which the GPU reads as absent."""
    b = r3d.Bricks(ctx, 256, 256, 256, 2, 2)
    page = np.full(64 + 9, 0xFFFFFFFF, np.uint32)
    page[:64] = 0
    page[0:4] = (2, 256, 256, 256)
    page[4:7] = (64, 2 | 2 << 10 | 2 << 20, 1)
    page[8:11] = (72, 1 | 1 << 10 | 1 << 20, 2)
    slots = np.zeros((4, 128 ** 3), np.uint8)
    for s, v in enumerate((200, 20, 50)):
        slots[s] = v
    page[64 + 4] = 0 | 200 << 24  # level 0 (0, 0, 1)
    page[64 + 1] = 1 | 20 << 24  # level 0 (1, 0, 0)
    page[64 + 5] = 0 | 200 << 24  # level 0 (1, 0, 1)
    page[72] = 2 | 50 << 24  # level 1
    b.write_page(page)
    b.write_slots(0, slots[:3])
    assert np.array_equal(b.read_page(), page)
    ob = orc.OracleBricks(page, np.concatenate([slots.reshape(-1), np.zeros(4 * 128 ** 3, np.uint8)]), 2)
    cam = scene.orbit_camera(10.0, 5.0, 1.8)
    lut = scene.tf_ramp_40_200()
    for thr in (0.0, 100.0 / 255.0):
        vs = [scene.frame_params(cam, 64, 48, mode=m, threshold=thr, lod_bias=-3.0, brick_mode=b.brick_mode)
              for m in (r3d.MODE_MIP, r3d.MODE_FULL)]
        for v in vs:
            check(ctx, b, ob, lut, [v], 64, 48)
    # slot 9 is past the 8 slots: absent on the GPU; the oracle's table marks it absent outright
    bad = page.copy()
    bad[64 + 2] = 9 | 90 << 24
    b.write_page(bad)
    v = scene.frame_params(cam, 64, 48, mode=r3d.MODE_MIP, lod_bias=-3.0, brick_mode=b.brick_mode)
    check(ctx, b, ob, lut, [v], 64, 48)
    b.close()


def test_rejections(ctx, bscene):
    b, _, _ = bscene
    cube = scene.frame_params(scene.orbit_camera(), 16, 16)
    with pytest.raises(r3d.R3DError):
        b.raycast([cube], 16, 16)  # a cube-mode view
    wrong = scene.frame_params(scene.orbit_camera(), 16, 16, brick_mode=(ABPA + 1) | 0x20000)
    with pytest.raises(r3d.R3DError):
        b.raycast([wrong], 16, 16)
    v = ctx.volume(32, 32, 32)
    v.upload(np.zeros(32 ** 3, np.uint8))
    with pytest.raises(r3d.R3DError):
        ctx.raycast(v, [scene.frame_params(scene.orbit_camera(), 16, 16, brick_mode=b.brick_mode)], 16, 16)
    small = r3d.Bricks(ctx, 300, 40, 40, 2, 1)  # 3 + 2 bricks, 1 slot
    with pytest.raises(r3d.R3DError):
        small.build(torch.zeros((40, 40, 300), dtype=torch.uint8, device="cuda"))
    for bad in ((0, 8, 8, 1, 1), (8, 8, 8, 0, 1), (8, 8, 8, 16, 1), (8, 8, 8, 1, 0), (8, 8, 8, 1, 256)):
        with pytest.raises(r3d.R3DError):
            r3d.Bricks(ctx, *bad)
