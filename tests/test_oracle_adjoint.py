"""The FP64 reverse-mode gradient of the oracle's differentiable raycast
(transfer table + density grid) pinned against FP64 central finite
differences, h = 1e-3, on 64 seeded entries. The reference has no
differentiable path; finite differences are what fixes the gradient."""
import numpy as np

from oracle import oracle as orc
from render3d_amd import scene
from render3d_amd._lib import MODE_FLAT, MODE_FULL


def gaussian_tf(centers=(90, 170), sigma=15.0):
    """Gaussian bumps in alpha (and a warm colour ramp) over the 256 entries."""
    x = np.arange(256, dtype=np.float64)
    a = np.zeros(256)
    for c in centers:
        a += np.exp(-0.5 * ((x - c) / sigma) ** 2)
    tf = np.zeros((256, 4))
    tf[:, 0] = np.clip(x / 255.0 * 1.1, 0, 1)
    tf[:, 1] = np.clip(x / 255.0 * 0.9, 0, 1)
    tf[:, 2] = np.clip(x / 255.0 * 0.7, 0, 1)
    tf[:, 3] = np.clip(a * 0.8, 0, 0.95)
    return tf


def small_problem(mode=MODE_FULL, n=32, g=6, W=20, H=16):
    vol = orc.synth_sheets(n, n, n, 2, 99)
    ov = orc.OracleVolume(vol, n, n, n)
    rng = np.random.default_rng(11)
    tf = gaussian_tf()
    grid = 1.0 + 0.3 * rng.uniform(-1, 1, (g, g, g))
    views = [scene.frame_params(scene.orbit_camera(40.0, 15.0, 1.8), W, H, mode=mode, step_voxels=0.7),
             scene.frame_params(scene.orbit_camera(160.0, -20.0, 1.6), W, H, mode=mode, step_voxels=0.7,
                                frame_index=3)]
    target = rng.uniform(0, 1, (len(views), H, W, 4))
    return ov, tf, grid, (g, g, g), views, target, W, H


def loss_and_grad(ov, tf, grid, gd, views, target, W, H, want_grad=True):
    L = 0.0
    dtf = np.zeros((256, 4))
    dg = np.zeros(grid.size)
    for i, v in enumerate(views):
        img, _, _ = orc.raycast_diff(ov, tf, grid, gd, [v], W, H, 0.98)
        r = img - target[i]
        L += float(np.sum(r * r))
        if want_grad:
            _, a, b = orc.raycast_diff(ov, tf, grid, gd, [v], W, H, 0.98, dimg=2 * r)
            dtf += a
            dg += b
    return L, dtf, dg


def check_fd(mode):
    ov, tf, grid, gd, views, target, W, H = small_problem(mode)
    L0, dtf, dg = loss_and_grad(ov, tf, grid, gd, views, target, W, H)
    rng = np.random.default_rng(5)
    # 32 TF entries where the forward sees data, 32 grid cells
    tf_cands = np.argwhere(np.abs(dtf) > 1e-9)
    tf_pick = tf_cands[rng.choice(len(tf_cands), 32, replace=False)]
    g_pick = rng.choice(grid.size, 32, replace=False)
    scale = max(np.abs(dtf).max(), np.abs(dg).max())
    checked, bad = 0, []

    def fd(setter, h):
        setter(+h)
        lp = loss_and_grad(ov, tf, grid, gd, views, target, W, H, False)[0]
        setter(-h)
        lm = loss_and_grad(ov, tf, grid, gd, views, target, W, H, False)[0]
        setter(0.0)
        return (lp - lm) / (2 * h)

    entries = [("tf", tuple(e)) for e in tf_pick] + [("grid", int(e)) for e in g_pick]
    for kind, idx in entries:
        if kind == "tf":
            base = tf[idx]

            def setter(d, idx=idx, base=base):
                tf[idx] = base + d
            adj = dtf[idx]
        else:
            flat = grid.reshape(-1)
            base = flat[idx]

            def setter(d, idx=idx, base=base, flat=flat):
                flat[idx] = base + d
            adj = dg[idx]
        f1 = fd(setter, 1e-3)
        f2 = fd(setter, 5e-4)
        if abs(f1 - f2) > 1e-5 * max(scale, 1.0) + 1e-3 * abs(f1):
            continue  # a 0.003 gate, the cutoff or a clamp sits within h: not differentiable here
        checked += 1
        rel = abs(f1 - adj) / max(abs(f1), abs(adj), 1e-6 * scale)
        if rel > 1e-3:
            bad.append((kind, idx, f1, adj, rel))
    return checked, bad


def test_adjoint_matches_fp64_finite_differences_full():
    checked, bad = check_fd(MODE_FULL)
    assert checked >= 48, checked
    assert not bad, bad


def test_adjoint_matches_fp64_finite_differences_flat():
    checked, bad = check_fd(MODE_FLAT)
    assert checked >= 48, checked
    assert not bad, bad


def test_diff_forward_matches_render_when_grid_is_one():
    """With g = 1 and an FP32 TF, the differentiable forward's colour is the
    plain renderer's (up to FP64 vs FP32 compositing) where no skip happens."""
    ov, tf, grid, gd, views, target, W, H = small_problem(MODE_FLAT)
    grid[:] = 1.0
    tf32 = tf.astype(np.float32)
    p = views[0]
    img, _, _ = orc.raycast_diff(ov, tf32.astype(np.float64), grid, gd, [p], W, H, 0.98)
    ref, _, _ = orc.raycast(ov, tf32, [p], W, H, alpha_cutoff=0.98)
    assert np.max(np.abs(img[..., :3] - ref[..., :3])) < 2e-4
