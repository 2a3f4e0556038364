"""The raycast oracle pinned without a runnable reference shader:
libm accuracy of the pinned math, the reference's own known answers
(tests/test_quick.c, tests/test_gpu.c), closed forms (empty volume, constant
volume composite, MIP = brute-force max) and exact integer preprocessing
(pyramid box filter, occupancy) against numpy brute force."""
import ctypes as C
import math

import numpy as np
import pytest

from oracle import oracle as orc
from render3d_amd import scene
from render3d_amd._lib import FrameParams, MODE_FLAT, MODE_FULL, MODE_MIP

f32 = np.float32


def ulps(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    sp = np.spacing(np.abs(b).astype(np.float32)).astype(np.float64)
    return np.abs(a - b) / sp


def test_log2_within_2_ulp():
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.uniform(1e-6, 1.0, 200000), rng.uniform(1.0, 4096.0, 200000),
                        np.float32(2.0) ** rng.integers(-140, 120, 2000).astype(np.float64)]).astype(np.float32)
    got = orc.math_batch(x, 0)
    ref = np.log2(x.astype(np.float64))
    big = np.abs(ref) > 1e-3
    assert np.max(ulps(got[big], ref[big])) <= 2.0
    assert np.max(np.abs(got[~big] - ref[~big])) < 1e-9 + 1e-7 * np.max(np.abs(ref[~big]), initial=0)
    assert orc.math_batch(np.array([0.0], np.float32), 0)[0] == -np.inf
    assert np.isnan(orc.math_batch(np.array([-1.0], np.float32), 0)[0])
    assert np.array_equal(orc.math_batch(np.array([1, 2, 4, 0.5, 1024], np.float32), 0), [0, 1, 2, -1, 10])


def test_exp2_within_2_ulp():
    x = np.linspace(-124.9, 127.4, 400001).astype(np.float32)
    got = orc.math_batch(x, 1)
    ref = np.exp2(x.astype(np.float64))
    assert np.max(ulps(got, ref)) <= 2.0
    assert np.array_equal(orc.math_batch(np.array([0, 1, -1, 10], np.float32), 1), [1, 2, 0.5, 1024])
    assert orc.math_batch(np.array([-200.0], np.float32), 1)[0] == 0.0
    assert orc.math_batch(np.array([200.0], np.float32), 1)[0] == np.inf


def test_pow_unit_and_pow24():
    b = np.linspace(0, 1, 10001).astype(np.float32)
    got = orc.math_batch(b, 2)  # b ** 1.7
    ref = np.power(b.astype(np.float64), np.float64(np.float32(1.7)))
    assert np.max(np.abs(got - ref)) < 4e-7
    p24 = orc.math_batch(b, 5)
    assert np.max(np.abs(p24 - b.astype(np.float64) ** 24)) < 1e-6


def test_sincos_turns():
    u = np.linspace(-2, 2, 100001).astype(np.float32)
    assert np.max(np.abs(orc.math_batch(u, 6) - np.sin(2 * np.pi * u.astype(np.float64)))) < 5e-7
    assert np.max(np.abs(orc.math_batch(u, 7) - np.cos(2 * np.pi * u.astype(np.float64)))) < 5e-7


def hash12_py(px, py, frame):
    """common.slang:57-61 in Python integers (independent restatement)."""
    m = 0xFFFFFFFF
    h = (px * 374761393 + py * 668265263 + frame * 2246822519) & m
    h = ((h ^ (h >> 13)) * 1274126177) & m
    return (h >> 8) / 16777216.0


def test_hash12_known_answers():
    L = orc.load()
    for px, py, fr in [(0, 0, 0), (1, 0, 0), (17, 33, 5), (1919, 1079, 123456), (65535, 65535, 4294967295)]:
        assert L.r3do_hash12(px, py, fr) == np.float32(hash12_py(px, py, fr))


# ---- scene helpers vs tests/test_quick.c known answers -------------------------

def feq(a, b):
    return abs(float(a) - float(b)) < 1e-6


def test_transfer_known_answers():  # test_quick.c:136-151
    lut = scene.tf_build(scene.tf_preset(0))
    assert lut[0, 0] == 0 and lut[0, 3] == 0
    assert lut[255, 0] == 255 and lut[255, 3] == 255
    assert lut[128, 0] == 128
    assert scene.tf_preset(99) is None and scene.TF_PRESET_COUNT == 3
    lut = scene.tf_build(scene.tf_preset(1))
    assert lut[20, 3] == 0 and lut[255, 3] == 255 and lut[100, 3] >= lut[60, 3]


def test_transfer_integer_interpolation():
    lut = scene.tf_ramp_40_200()
    assert lut[40, 3] == 0 and lut[200, 3] == 255
    assert lut[120, 3] == (0 * 80 + 255 * 80) // 160
    assert scene.tf_min_visible(lut) == 41
    assert scene.skip_gate(lut) == np.float32(np.float32(40.5) / np.float32(255))


def test_camera_known_answers():  # test_quick.c:99-134
    c = scene.Camera((0, 0, 0))
    r, u, fw = c.basis(2.0)
    assert feq(fw[0], 0) and feq(fw[1], 0) and feq(fw[2], 1)
    assert feq(np.dot(r, fw), 0) and feq(np.dot(u, fw), 0)
    assert feq(np.linalg.norm(u), math.tan(c.fov_y * 0.5))
    assert feq(np.linalg.norm(r), 2 * math.tan(c.fov_y * 0.5))
    c.look(0.0, 10.0)
    assert c.pitch < 1.5708
    c = scene.Camera((0, 0, 0))
    c.orbit_set((0.5, 0.5, 0.5), 2.0)
    assert feq(c.pos[0], 0.5) and feq(c.pos[1], 0.5) and feq(c.pos[2], -1.5)
    c.dist = f32(1.0)
    c.orbit_drag(3.14159265 / 2, 0.0)
    assert abs(c.pos[0] + 0.5) < 1e-5 and abs(c.pos[2] - 0.5) < 1e-5


def test_frame_params_layout_matches_oracle():
    assert C.sizeof(FrameParams) == 248
    offs = {name: getattr(FrameParams, name).offset for name, _ in FrameParams._fields_}
    # render_types.h:47-94 with float3 packed into 16-byte rows
    assert offs["step_voxels"] == 12 and offs["density"] == 28 and offs["max_mip"] == 60
    assert offs["viewport"] == 64 and offs["mode"] == 72 and offs["clip_orig"] == 120
    assert offs["vol_r0"] == 168 and offs["brick_mode"] == 216 and offs["overlay_flags"] == 244


# ---- preprocessing ---------------------------------------------------------------

def test_downsample_is_rounded_box():
    rng = np.random.default_rng(4)
    for dims in [(8, 8, 8), (9, 6, 5), (1, 4, 3)]:
        nx, ny, nz = dims
        v = rng.integers(0, 256, nx * ny * nz).astype(np.uint8)
        lv = orc.pyramid(v, nx, ny, nz, 2)
        (mx, my, mz), d = lv[1]
        a = v.reshape(nz, ny, nx).astype(np.int64)
        want = np.zeros((mz, my, mx), np.int64)
        for z in range(mz):
            for y in range(my):
                for x in range(mx):
                    s = 4
                    for dz in (0, 1):
                        for dy in (0, 1):
                            for dx in (0, 1):
                                s += a[min(2 * z + dz, nz - 1), min(2 * y + dy, ny - 1), min(2 * x + dx, nx - 1)]
                    want[z, y, x] = s // 8
        assert np.array_equal(d.reshape(mz, my, mx), want)


def test_occupancy_is_dilated_block_max():
    rng = np.random.default_rng(6)
    nx, ny, nz = 37, 20, 17
    v = (rng.random(nx * ny * nz) < 0.002).astype(np.uint8) * rng.integers(1, 256, nx * ny * nz).astype(np.uint8)
    (ox, oy, oz), occ = orc.occupancy(v, nx, ny, nz)
    a = v.reshape(nz, ny, nx)
    raw = np.zeros((oz, oy, ox), np.uint8)
    for z in range(oz):
        for y in range(oy):
            for x in range(ox):
                raw[z, y, x] = a[z * 8:z * 8 + 8, y * 8:y * 8 + 8, x * 8:x * 8 + 8].max()
    want = np.zeros_like(raw)
    for z in range(oz):
        for y in range(oy):
            for x in range(ox):
                want[z, y, x] = raw[max(z - 1, 0):z + 2, max(y - 1, 0):y + 2, max(x - 1, 0):x + 2].max()
    assert np.array_equal(occ.reshape(oz, oy, ox), want)


def test_synth_sheets_structure():
    n = 96
    v = orc.synth_sheets(n, n, n, 4, 5).reshape(n, n, n)
    # sheet k=2 sits at radius 38 +- 3 sin(2 pi z / 40); the background is ~20
    z = 10
    yy, xx = np.mgrid[0:n, 0:n]
    r = np.hypot(xx + 0.5 - n / 2, yy + 0.5 - n / 2) - 3 * math.sin(2 * math.pi * z / 40)
    on = np.abs(r - 38) < 0.7
    off = (np.abs(r - 33.5) < 0.5)
    assert v[z][on].mean() > 120 and v[z][off].mean() < 35
    assert 6 < v[z][off].std() < 12  # noise sigma 8 plus the +-5 background offsets


# ---- raycast closed forms ----------------------------------------------------------

def view(cam, w, h, **kw):
    return scene.frame_params(cam, w, h, **kw)


def test_empty_volume_is_background():
    n = 32
    ov = orc.OracleVolume(np.zeros(n ** 3, np.uint8), n, n, n)
    lut = scene.tf_unorm(scene.tf_build(scene.tf_preset(0)))
    img, cnt, _ = orc.raycast(ov, lut, [view(scene.orbit_camera(), 40, 30, step_voxels=0.5)], 40, 30)
    assert np.all(img == np.array([0.02, 0.02, 0.03, 1.0], np.float32))
    assert cnt["samples"] == cnt["skips"]


def _march_constant(ro, d, tn, tf, dt, jitter, alpha, cutoff):
    """FP64 closed form for a constant volume, lod pinned to 0."""
    t = tn + jitter * dt
    T, n = 1.0, 0
    while t < tf and n < 4096:
        n += 1
        T *= (1 - alpha)
        if 1 - T >= cutoff:
            break
        t += dt
    return 1 - T, n


def test_constant_volume_composite_closed_form():
    n, W, H = 32, 24, 20
    val = 150
    ov = orc.OracleVolume(np.full(n ** 3, val, np.uint8), n, n, n)
    lut8 = scene.tf_ramp_40_200()
    lut = scene.tf_unorm(lut8)
    cam = scene.orbit_camera(30.0, 15.0, 1.7)
    p = view(cam, W, H, mode=MODE_FLAT, step_voxels=0.5, density=0.3, lod_bias=-100.0)
    img, cnt, steps = orc.raycast(ov, lut, [p], W, H, alpha_cutoff=0.99, px_steps=True)
    # the TF is sampled at t = v*256 - 0.5 (256 texels over [0,1]), between texels
    v = val / 255.0
    tt = v * 256 - 0.5
    i0 = int(tt)
    w = tt - i0
    sa = float(lut[i0, 3]) * (1 - w) + float(lut[i0 + 1, 3]) * w
    a = 1 - (1 - sa) ** (0.3 * 0.5)
    col = float(lut[i0, 0]) * (1 - w) + float(lut[i0 + 1, 0]) * w
    right, up, fwd = (np.array(p.cam_right, np.float64), np.array(p.cam_up, np.float64),
                      np.array(p.cam_forward, np.float64))
    org = np.array(p.cam_origin, np.float64)
    checked = 0
    for py in range(H):
        for px in range(W):
            u, vv = (px + 0.5) / W, (py + 0.5) / H
            d = fwd + (u * 2 - 1) * right + (1 - vv * 2) * up
            d /= np.linalg.norm(d)
            with np.errstate(divide="ignore"):
                t0 = (0 - org) / d
                t1 = (1 - org) / d
            tn = max(np.max(np.minimum(t0, t1)), 0.0)
            tf = np.min(np.maximum(t0, t1))
            if tn >= tf:
                assert np.allclose(img[py, px, :3], [0.02, 0.02, 0.03])
                continue
            dt = 0.5 / n
            jit = hash12_py(px, py, 0)
            frac = ((tf - tn) / dt - jit) % 1.0
            if min(frac, 1 - frac) < 1e-3:
                continue  # the last sample sits on the exit face: count is rounding-sensitive
            acc, ns = _march_constant(org, d, tn, tf, dt, jit, a, 0.99)
            assert steps[py, px] == ns, (px, py)
            want = col * acc + (1 - acc) * 0.02
            assert abs(img[py, px, 0] - want) < 2e-5, (px, py, img[py, px, 0], want)
            checked += 1
    assert checked > W * H // 4


def test_mip_is_bruteforce_max():
    n, W, H = 24, 16, 12
    rng = np.random.default_rng(8)
    vol = rng.integers(1, 230, n ** 3).astype(np.uint8)
    ov = orc.OracleVolume(vol, n, n, n)
    lut = scene.tf_unorm(scene.tf_ramp_40_200())
    p = view(scene.orbit_camera(20.0, 10.0, 1.8), W, H, mode=MODE_MIP, step_voxels=0.7, lod_bias=-100.0)
    img, _, _ = orc.raycast(ov, lut, [p], W, H)
    a = vol.reshape(n, n, n).astype(np.float64)

    def tri(q):
        t = np.clip(q * n - 0.5, 0, n - 1)
        i0 = t.astype(int)
        i1 = np.minimum(i0 + 1, n - 1)
        f = t - i0
        c = 0.0
        for dz in (0, 1):
            for dy in (0, 1):
                for dx in (0, 1):
                    w = (f[0] if dx else 1 - f[0]) * (f[1] if dy else 1 - f[1]) * (f[2] if dz else 1 - f[2])
                    c += w * a[(i1 if dz else i0)[2], (i1 if dy else i0)[1], (i1 if dx else i0)[0]]
        return c / 255.0

    right, up, fwd = (np.array(p.cam_right, np.float64), np.array(p.cam_up, np.float64),
                      np.array(p.cam_forward, np.float64))
    org = np.array(p.cam_origin, np.float64)
    for py in range(H):
        for px in range(W):
            u, vv = (px + 0.5) / W, (py + 0.5) / H
            d = fwd + (u * 2 - 1) * right + (1 - vv * 2) * up
            d /= np.linalg.norm(d)
            with np.errstate(divide="ignore"):
                t0, t1 = (0 - org) / d, (1 - org) / d
            tn, tf = max(np.max(np.minimum(t0, t1)), 0.0), np.min(np.maximum(t0, t1))
            if tn >= tf:
                continue
            dt = 0.7 / n
            t = tn + hash12_py(px, py, 0) * dt
            m = 0.0
            while t < tf - 1e-6:
                m = max(m, tri(org + t * d))
                t += dt
            assert abs(img[py, px, 0] - m) < 2e-5, (px, py)


def gen_shells(n):
    """tests/test_gpu.c:17-22 (the reference's GPU conformance volume)."""
    inv = 2.0 / (n - 1)
    z, y, x = np.mgrid[0:n, 0:n, 0:n].astype(np.float32)
    X, Y, Z = x * np.float32(inv) - 1, y * np.float32(inv) - 1, z * np.float32(inv) - 1
    r = np.sqrt(X * X + Y * Y + Z * Z)
    s = 0.5 + 0.5 * np.sin(r * 40.0)
    fade = (r < 0.9).astype(np.float32)
    return (255.0 * s * s * fade).astype(np.uint8)


def test_against_reference_gpu_conformance_march():
    """tests/test_gpu.c:97-157 restated in numpy: MODE_FLAT, identity ramp,
    camera at (0.5, 0.5, -1.5), every 3rd pixel, tolerance mean <= 2 LSB and
    max <= 64 LSB as the reference asserts (its CPU march has no LOD and no
    empty-space skip; lod_bias pins lod 0 here)."""
    n, W, H = 64, 96, 72
    vol = gen_shells(n)
    ov = orc.OracleVolume(vol.reshape(-1), n, n, n)
    lut = scene.tf_unorm(scene.tf_build(scene.tf_preset(0)))
    cam = scene.Camera((0.5, 0.5, -1.5))
    p = view(cam, W, H, mode=MODE_FLAT, step_voxels=1.0, lod_bias=-100.0)
    img, _, _ = orc.raycast(ov, lut, [p], W, H, alpha_cutoff=0.98)
    right, up, fwd = cam.basis(f32(W) / f32(H))
    a = vol.astype(np.float32) / np.float32(255)
    errs = []
    for py in range(0, H, 3):
        for px in range(0, W, 3):
            u, vv = (px + 0.5) / W, (py + 0.5) / H
            d = fwd + np.float32(u * 2 - 1) * right + np.float32(1 - vv * 2) * up
            d = d / np.linalg.norm(d)
            o = cam.pos.astype(np.float64)
            tn, tf = -1e30, 1e30
            for ax in range(3):
                t0, t1 = (0 - o[ax]) / d[ax], (1 - o[ax]) / d[ax]
                tn, tf = max(tn, min(t0, t1)), min(tf, max(t0, t1))
            tn = max(tn, 0.0)
            accr = acca = 0.0
            if tn < tf:
                dt = 1.0 / n
                t = tn + hash12_py(px, py, 0) * dt
                k = 0
                while t < tf and k < 4096:
                    q = o + t * d
                    tt = np.clip(q * n - 0.5, 0, n - 1)
                    i0 = tt.astype(int)
                    i1 = np.minimum(i0 + 1, n - 1)
                    f = tt - i0
                    c = 0.0
                    for dz in (0, 1):
                        for dy in (0, 1):
                            for dx in (0, 1):
                                w = (f[0] if dx else 1 - f[0]) * (f[1] if dy else 1 - f[1]) * (f[2] if dz else 1 - f[2])
                                c += w * a[(i1 if dz else i0)[2], (i1 if dy else i0)[1], (i1 if dx else i0)[0]]
                    tt2 = min(max(c * 256 - 0.5, 0), 255)
                    j0 = int(tt2)
                    j1 = min(j0 + 1, 255)
                    al = j0 / 255 + (j1 / 255 - j0 / 255) * (tt2 - j0)
                    if al > 0.003:
                        accr += (1 - acca) * al * al
                        acca += (1 - acca) * al
                        if acca >= 0.98:
                            break
                    k += 1
                    t += dt
            ref = accr + (1 - acca) * 0.02
            errs.append(abs(ref - img[py, px, 0]) * 255)
    errs = np.array(errs)
    assert errs.mean() <= 2.0 and errs.max() <= 64.0, (errs.mean(), errs.max())


def test_step_invariance_of_opacity_correction():
    """raycast.slang:905-906: 1 - (1-a)^(density*step) keeps appearance
    invariant under step changes (constant volume, lod pinned)."""
    n, W, H = 32, 16, 12
    ov = orc.OracleVolume(np.full(n ** 3, 90, np.uint8), n, n, n)
    lut = scene.tf_unorm(scene.tf_ramp_40_200())
    cam = scene.orbit_camera(10.0, 5.0, 1.8)
    outs = []
    for step in (0.25, 0.5, 1.0):
        p = view(cam, W, H, mode=MODE_FLAT, step_voxels=step, density=0.05, lod_bias=-100.0)
        img, _, _ = orc.raycast(ov, lut, [p], W, H, alpha_cutoff=0.999)
        outs.append(img[..., 0])
    assert np.max(np.abs(outs[0] - outs[1])) < 0.03 and np.max(np.abs(outs[1] - outs[2])) < 0.05


def test_unsupported_view_raises():
    n = 16
    ov = orc.OracleVolume(np.zeros(n ** 3, np.uint8), n, n, n)
    p = view(scene.orbit_camera(), 8, 8)
    p.brick_mode = 3
    with pytest.raises(ValueError):
        orc.raycast(ov, scene.tf_unorm(scene.tf_ramp_40_200()), [p], 8, 8)


def test_lod_policy_known_answers():  # tests/test_quick.c:64-71
    assert scene.lod_pick(0.0, 0.001, 8192.0, 8) == 0
    assert scene.lod_pick(1.0, 0.001, 8192.0, 8) == 3  # 8.2 voxels per pixel
    assert scene.lod_pick(100.0, 0.01, 8192.0, 8) == 7  # clamped to the coarsest
    assert [scene.lod_fallback(2, i, 5) for i in range(5)] == [2, 3, 4, 1, 0]
    assert scene.lod_fallback(2, 5, 5) == scene.LOD_NONE
    # every level visited once; the pick at exact powers of two
    for n in range(1, 7):
        for d in range(n):
            assert sorted(scene.lod_fallback(d, i, n) for i in range(n)) == list(range(n))
    assert [scene.lod_pick(d, 1.0, 1.0, 12) for d in (1.0, 2.0, 3.99, 4.0, 1024.0)] == [0, 1, 1, 2, 10]
