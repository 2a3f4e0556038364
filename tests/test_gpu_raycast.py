"""HIP raycaster, pyramid, occupancy and generator parity against the oracle
(host transcription of raycast.slang cube mode), through the C ABI. The
raycast bar is max-abs <= 1e-5 in FP32 and identical per-pixel sample counts."""
import math

import numpy as np
import pytest

import render3d_amd as r3d
from oracle import oracle as orc
from render3d_amd import scene

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

TOL = 1e-5


def test_synth_matches_host_generator(ctx):
    for (nx, ny, nz, ns, seed) in [(64, 64, 64, 3, 9), (96, 80, 48, 4, 123)]:
        g = ctx.synth_sheets(nx, ny, nz, ns, seed).cpu().numpy()
        h = orc.synth_sheets(nx, ny, nz, ns, seed)
        assert np.array_equal(g, h), (nx, ny, nz, int(np.sum(g != h)))


@pytest.mark.parametrize("dims", [(64, 64, 64), (96, 80, 40), (33, 17, 70)])
def test_pyramid_and_occupancy_exact(ctx, dims):
    nx, ny, nz = dims
    vol = orc.synth_sheets(nx, ny, nz, 3, 77)
    v = ctx.volume(nx, ny, nz)
    v.upload(vol)
    ref = orc.pyramid(vol, nx, ny, nz, v.nlevels)
    for l in range(v.nlevels):
        assert v.level_dims[l] == ref[l][0]
        assert np.array_equal(v.read_level(l), ref[l][1]), f"level {l}"
    odim, occ = orc.occupancy(vol, nx, ny, nz)
    assert v.odim == odim
    assert np.array_equal(v.read_occupancy(), occ)


def test_decode_into_bricks_equals_upload(ctx):
    n = 64
    vol = ctx.synth_sheets(n, n, n, 3, 21)
    step = torch.from_numpy(orc.step_linear()).cuda()
    levels, hdr = ctx.encode_chunks(vol, n, n, n, step)
    flat = ctx.decode_chunks(headers=hdr, step_tables=step, levels=levels)
    v = ctx.volume(n, n, n, 4)
    offs, base, sy, sz = v.chunk_targets()
    ctx.decode(ctx.make_batch(headers=hdr, step_tables=step, out=base, levels=levels, out_offsets=offs,
                              stride_y=sy, stride_z=sz))
    v.build()
    lin = flat.cpu().numpy().reshape(n // 16, n // 16, n // 16, 16, 16, 16).transpose(0, 3, 1, 4, 2, 5).reshape(-1)
    w = ctx.volume(n, n, n, 4)
    w.upload(lin)
    for l in range(4):
        assert np.array_equal(v.read_level(l), w.read_level(l))
    assert np.array_equal(v.read_occupancy(), w.read_occupancy())


@pytest.mark.parametrize("n,nlev", [(64, 4), (48, 0), (32, 1)])
def test_volume_decode_and_build_from_chunks(ctx, n, nlev):
    vol = ctx.synth_sheets(n, n, n, 3, 5 + n)
    step = torch.from_numpy(orc.step_linear()).cuda()
    levels, hdr = ctx.encode_chunks(vol, n, n, n, step)
    flat = ctx.decode_chunks(headers=hdr, step_tables=step, levels=levels)
    lin = flat.cpu().numpy().reshape(n // 16, n // 16, n // 16, 16, 16, 16).transpose(0, 3, 1, 4, 2, 5).reshape(-1)
    ref = orc.pyramid(lin, n, n, n, orc.full_levels(n, n, n) if nlev == 0 else nlev)
    odim, occ = orc.occupancy(lin, n, n, n)
    a = ctx.volume(n, n, n, nlev)
    a.decode(ctx.make_batch(headers=hdr, step_tables=step, out=0, levels=levels))
    b = ctx.volume(n, n, n, nlev)
    b.build_from_chunks(flat)
    for v in (a, b):
        assert v.nlevels == len(ref)
        for l in range(v.nlevels):
            assert np.array_equal(v.read_level(l), ref[l][1]), l
        assert np.array_equal(v.read_occupancy(), occ)


@pytest.mark.parametrize("dims", [(128, 48, 80), (64, 16, 32), (192, 32, 16), (80, 48, 32)])
def test_build_from_chunks_shapes(ctx, dims):
    """Level 0..n and the occupancy grid from flat chunks of non-cubic volumes
    against the oracle's pyramid: rows of 8, 4 and 12 bricks (taken four at a
    time), a single brick row or plane (the y and z aprons clamp to the
    brick's own last layer), and a row of 5 (one brick at a time)."""
    nx, ny, nz = dims
    bx, by, bz = nx // 16, ny // 16, nz // 16
    rng = np.random.default_rng(nx * 7 + ny * 3 + nz)
    lin = orc.synth_sheets(nx, ny, nz, 4, 31 + nx)
    lin[rng.random(lin.shape) < 0.02] = 255  # isolated bright voxels: every octant and apron byte matters
    chunks = lin.reshape(bz, 16, by, 16, bx, 16).transpose(0, 2, 4, 1, 3, 5).reshape(-1, 4096).copy()
    ref = orc.pyramid(lin, nx, ny, nz, orc.full_levels(nx, ny, nz))
    odim, occ = orc.occupancy(lin, nx, ny, nz)
    v = ctx.volume(nx, ny, nz, 0)
    v.build_from_chunks(torch.from_numpy(chunks).cuda())
    assert v.nlevels == len(ref)
    for l in range(v.nlevels):
        assert np.array_equal(v.read_level(l), ref[l][1]), l
    assert v.odim == odim
    assert np.array_equal(v.read_occupancy(), occ)


def _views(width, height):
    lut = scene.tf_ramp_40_200()
    gate = scene.skip_gate(lut)
    cam = scene.orbit_camera(45.0, 10.0, 2.0)
    base = dict(step_voxels=0.5, gate=gate)
    out = {
        "full": [scene.frame_params(cam, width, height, **base)],
        "flat": [scene.frame_params(cam, width, height, mode=r3d.MODE_FLAT, **base)],
        "mip": [scene.frame_params(cam, width, height, mode=r3d.MODE_MIP, **base)],
        "depth": [scene.frame_params(cam, width, height, mode=r3d.MODE_DEPTH, **base)],
        "heatmap": [scene.frame_params(cam, width, height, mode=r3d.MODE_HEATMAP, **base)],
        "raydir": [scene.frame_params(cam, width, height, mode=r3d.MODE_RAYDIR, **base)],
        "lodbias": [scene.frame_params(cam, width, height, lod_bias=1.5, frame_index=7, **base)],
        "threshold": [scene.frame_params(cam, width, height, threshold=0.3, **base)],
        "half": [scene.frame_params(cam, width, height, view_flags=r3d.VIEW_HALF, **base)],
    }
    ortho = scene.frame_params(cam, width, height, view_flags=r3d.VIEW_ORTHO, **base)
    ortho.cam_right[:] = [x * 0.9 for x in ortho.cam_right]
    out["ortho"] = [ortho]
    rot = scene.frame_params(scene.orbit_camera(-30.0, 25.0, 1.6), width, height, step_voxels=1.0, gate=gate)
    a = math.radians(20)
    rot.vol_r0[:] = [math.cos(a), 0.0, math.sin(a)]
    rot.vol_r1[:] = [0.0, 1.0, 0.0]
    rot.vol_r2[:] = [-math.sin(a), 0.0, math.cos(a)]
    rot.vol_tx, rot.vol_ty, rot.vol_tz = 0.05, -0.02, 0.01
    out["rotated"] = [rot]
    inside = scene.Camera((0.45, 0.55, 0.3))
    inside.yaw = np.float32(0.3)
    out["inside"] = [scene.frame_params(inside, width, height, step_voxels=0.75, gate=gate)]
    # several views of mixed modes and sizes in one call
    w2, h2 = width // 2, height // 2
    out["multi"] = [
        scene.frame_params(cam, w2, h2, view_org=(0, 0), **base),
        scene.frame_params(scene.orbit_camera(120.0, -20.0, 2.2), w2, h2, mode=r3d.MODE_MIP, view_org=(w2, 0),
                           **base),
        scene.frame_params(scene.orbit_camera(200.0, 5.0, 1.8), w2, h2, mode=r3d.MODE_FLAT, view_org=(0, h2),
                           **base),
        scene.frame_params(scene.orbit_camera(300.0, 40.0, 2.0), w2 - 3, h2 - 5, view_org=(w2, h2), **base),
    ]
    return lut, out


@pytest.fixture(scope="module")
def scene96(ctx):
    dims = (96, 80, 72)
    vol = orc.synth_sheets(*dims, 4, 31)
    v = ctx.volume(*dims)
    v.upload(vol)
    ov = orc.OracleVolume(vol, *dims)
    return v, ov


@pytest.mark.parametrize("case", ["full", "flat", "mip", "depth", "heatmap", "raydir", "lodbias", "threshold",
                                  "half", "ortho", "rotated", "inside", "multi"])
def test_raycast_matches_oracle(ctx, scene96, case):
    v, ov = scene96
    W, H = 120, 88
    lut, views = _views(W, H)
    vs = views[case]
    ctx.set_transfer(lut)
    out = torch.zeros((H, W, 4), dtype=torch.float32, device="cuda")
    res = ctx.raycast(v, vs, W, H, alpha_cutoff=0.99, out=out, px_steps=True, stats=True, rgba8=True)
    img = res["rgba"].cpu().numpy()
    ref, cnt, steps = orc.raycast(ov, scene.tf_unorm(lut), vs, W, H, alpha_cutoff=0.99, px_steps=True)
    err = float(np.max(np.abs(img - ref)))
    print(f"{case}: max abs {err:g}, samples {cnt['samples']}")
    assert err <= TOL, f"{case}: max abs {err}"
    assert np.array_equal(res["px_steps"].cpu().numpy().astype(np.uint32), steps), case
    st = res["stats"]
    assert (st["samples"], st["skips"], st["shaded"], st["rays"]) == \
        (cnt["samples"], cnt["skips"], cnt["shaded"], cnt["rays"]), (case, st, cnt)
    want8 = (np.clip(ref, 0, 1) * np.float32(255) + np.float32(0.5)).astype(np.uint8)
    want8[ref[..., 3] == 0] = 0  # pixels no view covers are left untouched (zero)
    assert np.array_equal(res["rgba8"].cpu().numpy(), want8)


@pytest.mark.parametrize("case", ["full", "lodbias", "half", "ortho", "rotated", "inside", "multi"])
def test_raycast_fast_quality_matches_oracle(ctx, scene96, case):
    """R3D_QUALITY_FAST (r3d_set_quality): FULL-mode views take the four
    tetrahedral gradient taps (raycast.slang:919-955); the other modes of a
    mixed call are unchanged."""
    v, ov = scene96
    W, H = 120, 88
    lut, views = _views(W, H)
    vs = views[case]
    ctx.set_transfer(lut)
    ctx.set_quality(r3d.QUALITY_FAST)
    try:
        out = torch.zeros((H, W, 4), dtype=torch.float32, device="cuda")  # pixels no view covers stay 0
        res = ctx.raycast(v, vs, W, H, alpha_cutoff=0.99, out=out, px_steps=True, stats=True)
    finally:
        ctx.set_quality(r3d.QUALITY_FULL)
    img = res["rgba"].cpu().numpy()
    ref, cnt, steps = orc.raycast(ov, scene.tf_unorm(lut), vs, W, H, alpha_cutoff=0.99, px_steps=True,
                                  quality=orc.QUALITY_FAST)
    assert float(np.max(np.abs(img - ref))) <= TOL
    assert np.array_equal(res["px_steps"].cpu().numpy().astype(np.uint32), steps)
    st = res["stats"]
    assert (st["samples"], st["skips"], st["shaded"], st["rays"]) == \
        (cnt["samples"], cnt["skips"], cnt["shaded"], cnt["rays"])
    assert cnt["shaded"] > 0
    full, _, _ = orc.raycast(ov, scene.tf_unorm(lut), vs, W, H, alpha_cutoff=0.99)
    assert np.max(np.abs(full - ref)) > 1e-3  # the four taps do shade differently


@pytest.fixture(scope="module")
def sparse96(ctx):
    """Two sheets (radii 20 and 29 about the z axis) with everything below 90
    cleared: sheets on an empty background, so the dilated 8^3 occupancy is
    zero beyond the outer sheet (15% of the blocks) and rays take the
    empty-space jump (raycast.slang:829-869) before and after their hits."""
    dims = (96, 80, 72)
    vol = orc.synth_sheets(*dims, 2, 31)
    vol[vol < 90] = 0
    v = ctx.volume(*dims)
    v.upload(vol)
    return v, orc.OracleVolume(vol, *dims)


@pytest.mark.parametrize("case", ["full", "fast", "flat", "mip", "depth", "heatmap", "lodbias", "ortho", "rotated",
                                  "multi"])
def test_raycast_with_skips_matches_oracle(ctx, sparse96, case):
    v, ov = sparse96
    W, H = 120, 88
    lut, views = _views(W, H)
    vs = views["full" if case == "fast" else case]
    quality = r3d.QUALITY_FAST if case == "fast" else r3d.QUALITY_FULL
    ctx.set_transfer(lut)
    ctx.set_quality(quality)
    try:
        out = torch.zeros((H, W, 4), dtype=torch.float32, device="cuda")  # pixels no view covers stay 0
        res = ctx.raycast(v, vs, W, H, alpha_cutoff=0.99, out=out, px_steps=True, stats=True)
    finally:
        ctx.set_quality(r3d.QUALITY_FULL)
    img = res["rgba"].cpu().numpy()
    ref, cnt, steps = orc.raycast(ov, scene.tf_unorm(lut), vs, W, H, alpha_cutoff=0.99, px_steps=True,
                                  quality=quality)
    err = float(np.max(np.abs(img - ref)))
    print(f"{case}: max abs {err:g}, samples {cnt['samples']}, skips {cnt['skips']}, shaded {cnt['shaded']}")
    assert cnt["skips"] > cnt["rays"] // 4, cnt  # the jump is taken, and not only at the faces
    assert err <= TOL
    assert np.array_equal(res["px_steps"].cpu().numpy().astype(np.uint32), steps)
    st = res["stats"]
    assert (st["samples"], st["skips"], st["shaded"], st["rays"]) == \
        (cnt["samples"], cnt["skips"], cnt["shaded"], cnt["rays"])
    if case in ("full", "fast", "lodbias", "rotated", "multi"):
        assert cnt["shaded"] > 0


@pytest.mark.parametrize("case", ["full", "mip", "flat", "lodbias", "half", "multi"])
def test_measurement_mode_matches_oracle(ctx, scene96, case):
    """R3D_RAYCAST_MEASURE: the measurement instances render the production
    kernel's image bit for bit, count blended samples as the oracle does, and
    mark exactly the bricks whose footprints the oracle reads."""
    v, ov = scene96
    W, H = 120, 88
    lut, views = _views(W, H)
    vs = views[case]
    ctx.set_transfer(lut)
    prod = ctx.raycast(v, vs, W, H, alpha_cutoff=0.99, out=torch.zeros((H, W, 4), device="cuda"), stats=True)
    touched = torch.zeros((v.brick_count + 31) // 32, dtype=torch.int32, device="cuda")
    counters = torch.zeros(6, dtype=torch.int64, device="cuda")
    res = ctx.raycast(v, vs, W, H, alpha_cutoff=0.99, out=torch.zeros((H, W, 4), device="cuda"), stats=True,
                      measure=True, touched=touched)
    ctx.raycast(v, vs, W, H, alpha_cutoff=0.99, counters=counters, measure=True)
    assert torch.equal(res["rgba"], prod["rgba"])
    bits = ov.track_touched()
    try:
        _, cnt, _ = orc.raycast(ov, scene.tf_unorm(lut), vs, W, H, alpha_cutoff=0.99)
    finally:
        ov.track_touched(False)
    st = res["stats"]
    want = [cnt[k] for k in ("samples", "skips", "shaded", "rays", "blended", "blended_shaded")]
    assert [st[k] for k in ("samples", "skips", "shaded", "rays", "blended", "blended_shaded")] == want
    assert counters.cpu().numpy().tolist() == want
    if case in ("full", "lodbias", "multi"):
        assert cnt["blended"] > 0 and cnt["blended_shaded"] > 0
    got = touched.cpu().numpy().view(np.uint32)
    assert np.array_equal(got, bits), (int(np.unpackbits(got.view(np.uint8)).sum()),
                                       int(np.unpackbits(bits.view(np.uint8)).sum()))


def test_raycast_f32_tf_and_empty_volume(ctx):
    n = 48
    v = ctx.volume(n, n, n)
    v.upload(np.zeros(n ** 3, np.uint8))
    lut = scene.tf_unorm(scene.tf_build(scene.tf_preset(1)))
    tf = torch.from_numpy(lut.reshape(-1)).cuda()
    p = scene.frame_params(scene.orbit_camera(30.0, 10.0, 2.0), 64, 48, step_voxels=0.5, gate=0.0)
    res = ctx.raycast(v, [p], 64, 48, tf=tf, stats=True)
    img = res["rgba"].cpu().numpy()
    bg = np.array([0.02, 0.02, 0.03, 1.0], np.float32)
    assert np.all(img == bg)
    assert res["stats"]["samples"] == res["stats"]["skips"]


def test_unsupported_views_rejected(ctx, scene96):
    v, _ = scene96
    p = scene.frame_params(scene.orbit_camera(), 32, 32)
    p.slab_grid = 1
    with pytest.raises(r3d.R3DError):
        ctx.raycast(v, [p], 32, 32)
    p.slab_grid = 0
    p.mode = 9
    with pytest.raises(r3d.R3DError):
        ctx.raycast(v, [p], 32, 32)


@pytest.mark.parametrize("dims", [(128, 64, 96), (64, 48, 160), (64, 32, 16), (48, 64, 64)])
def test_volume_decode_equals_decode_then_build(ctx, dims):
    """r3d_hip_volume_decode (the chunk batch decoded into the volume's
    staging, then built) on shapes with and without the four-brick build
    groups: every level and the occupancy equal a separate decode + build of
    the same chunks, dense and sparse, twice in a row on one volume."""
    nx, ny, nz = dims
    vol = ctx.synth_sheets(nx, ny, nz, 3, 17)
    step = torch.from_numpy(orc.step_linear()).cuda()
    levels, hdr = ctx.encode_chunks(vol, nx, ny, nz, step)
    flat = ctx.decode_chunks(headers=hdr, step_tables=step, levels=levels)
    ref = ctx.volume(nx, ny, nz, 3)
    ref.build_from_chunks(flat)
    lv = levels.cpu().numpy()
    cnt = (lv != 0).sum(1)
    offs = np.zeros(lv.shape[0] + 1, np.uint32)
    offs[1:] = np.cumsum(cnt)
    pos = np.nonzero(lv.reshape(-1))[0]
    pairs = ((pos % 4096) | (lv.reshape(-1)[pos].view(np.uint16).astype(np.uint32) << 16)).astype(np.uint32)
    batches = [ctx.make_batch(headers=hdr, step_tables=step, out=0, levels=levels),
               ctx.make_batch(headers=hdr, step_tables=step, out=0, pairs=torch.from_numpy(pairs).cuda(),
                              pair_offsets=torch.from_numpy(offs).cuda())]
    for bt in batches:
        v = ctx.volume(nx, ny, nz, 3)
        for _ in range(2):  # twice: the second run reuses the staging
            v.decode(bt)
        torch.cuda.synchronize()
        for l in range(3):
            assert np.array_equal(v.read_level(l), ref.read_level(l)), (dims, l)
        assert np.array_equal(v.read_occupancy(), ref.read_occupancy()), dims
