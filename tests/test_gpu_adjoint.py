"""HIP differentiable raycast and adjoint against the FP64 oracle: forward
image, gradient (relative error <= 1e-3) and FP64 central finite differences;
then the torch.autograd harness and a short Adam fit."""
import numpy as np
import pytest

from oracle import oracle as orc
from render3d_amd import scene
from render3d_amd._lib import MODE_FLAT, MODE_FULL

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from test_oracle_adjoint import gaussian_tf, small_problem  # noqa: E402


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def gpu_setup(ctx, ov_dims, seed_vol, views, W, H):
    from render3d_amd.autograd import DiffRenderer
    n = ov_dims
    vol = orc.synth_sheets(n, n, n, 2, seed_vol)
    v = ctx.volume(n, n, n)
    v.upload(vol)
    return v, DiffRenderer(ctx, v, views, W, H, alpha_cutoff=0.98)


@pytest.mark.parametrize("mode", [MODE_FULL, MODE_FLAT])
def test_adjoint_matches_fp64_reference(ctx, mode):
    ov, tf, grid, gd, views, target, W, H = small_problem(mode)
    tf32 = tf.astype(np.float32)
    g32 = grid.astype(np.float32)
    dtf_ref = np.zeros((256, 4))
    dg_ref = np.zeros(grid.size)
    dtf_gpu = np.zeros((256, 4))
    dg_gpu = np.zeros(grid.size)
    for i, vw in enumerate(views):
        v, R = gpu_setup(ctx, 32, 99, [vw], W, H)
        t_tf = torch.from_numpy(tf32).cuda()
        t_g = torch.from_numpy(g32).cuda()
        img = R.forward_raw(t_tf, t_g)
        torch.cuda.synchronize()
        ref, _, _ = orc.raycast_diff(ov, tf32.astype(np.float64), g32.astype(np.float64), gd, [vw], W, H, 0.98)
        err = np.abs(img.cpu().numpy() - ref)
        # FP32 vs FP64: identical except where a sample's alpha sits on the 0.003 gate
        assert np.mean(err < 1e-4) > 0.999 and err.max() < 1e-2, (np.mean(err < 1e-4), err.max())
        dimg = 2 * (ref - target[i])
        a, b = R.adjoint_raw(t_tf, t_g, torch.from_numpy(dimg.astype(np.float32)).cuda())
        torch.cuda.synchronize()
        dtf_gpu += a.cpu().numpy()
        dg_gpu += b.cpu().numpy().reshape(-1)
        _, ra, rb = orc.raycast_diff(ov, tf32.astype(np.float64), g32.astype(np.float64), gd, [vw], W, H, 0.98,
                                     dimg=dimg)
        dtf_ref += ra
        dg_ref += rb
    assert rel(dtf_gpu, dtf_ref) <= 1e-3, rel(dtf_gpu, dtf_ref)
    assert rel(dg_gpu, dg_ref) <= 1e-3, rel(dg_gpu, dg_ref)


def test_autograd_harness_and_finite_differences(ctx):
    ov, tf, grid, gd, views, target, W, H = small_problem(MODE_FULL)
    v, R = gpu_setup(ctx, 32, 99, views[:1], W, H)
    t_tf = torch.from_numpy(tf.astype(np.float32)).cuda().requires_grad_(True)
    t_g = torch.from_numpy(grid.astype(np.float32)).cuda().requires_grad_(True)
    tgt = torch.from_numpy(target[0].astype(np.float32)).cuda()
    loss = ((R(t_tf, t_g) - tgt) ** 2).sum()
    loss.backward()
    gtf = t_tf.grad.cpu().numpy().astype(np.float64)
    gg = t_g.grad.cpu().numpy().reshape(-1).astype(np.float64)
    # FP64 central differences of the same loss through the FP64 forward
    tfd = tf.astype(np.float32).astype(np.float64)
    gdd = grid.astype(np.float32).astype(np.float64)

    def L():
        img, _, _ = orc.raycast_diff(ov, tfd, gdd, gd, views[:1], W, H, 0.98)
        return float(np.sum((img - target[0].astype(np.float32)) ** 2))

    rng = np.random.default_rng(9)
    cand = np.argwhere(np.abs(gtf) > 1e-6)
    picks = [("tf", tuple(c)) for c in cand[rng.choice(len(cand), 32, replace=False)]]
    picks += [("g", int(i)) for i in rng.choice(gdd.size, 32, replace=False)]
    scale = max(np.abs(gtf).max(), np.abs(gg).max())
    checked = 0
    for kind, idx in picks:
        arr = tfd if kind == "tf" else gdd.reshape(-1)
        base = arr[idx]
        vals = []
        for h in (1e-3, 5e-4):
            arr[idx] = base + h
            lp = L()
            arr[idx] = base - h
            lm = L()
            arr[idx] = base
            vals.append((lp - lm) / (2 * h))
        if abs(vals[0] - vals[1]) > 1e-5 * max(scale, 1.0) + 1e-3 * abs(vals[0]):
            continue
        got = gtf[idx] if kind == "tf" else gg[idx]
        checked += 1
        assert abs(got - vals[0]) <= 1e-3 * max(abs(vals[0]), abs(got)) + 1e-5 * scale, (kind, idx, got, vals[0])
    assert checked >= 40


def test_adam_fit_lowers_the_loss(ctx):
    from render3d_amd.autograd import DiffRenderer
    n, W, H = 48, 32, 32
    vol = ctx.synth_sheets(n, n, n, 3, 4242)
    v = ctx.volume(n, n, n)
    v.upload(vol)
    views = [scene.frame_params(scene.orbit_camera(30.0 + 90 * i, 10.0, 1.8), W, H, step_voxels=0.5)
             for i in range(4)]
    gt = torch.from_numpy(gaussian_tf().astype(np.float32)).cuda()
    grid1 = torch.ones((8, 8, 8), dtype=torch.float32, device="cuda")
    Rs = [DiffRenderer(ctx, v, [vw], W, H) for vw in views]
    targets = [R.forward_raw(gt, grid1).clone() for R in Rs]
    tf = torch.full((256, 4), 0.3, dtype=torch.float32, device="cuda", requires_grad=True)
    grid = torch.ones((8, 8, 8), dtype=torch.float32, device="cuda", requires_grad=True)
    opt = torch.optim.Adam([tf, grid], lr=0.02)
    losses = []
    for _ in range(40):
        opt.zero_grad()
        loss = sum(((R(tf, grid) - t) ** 2).sum() for R, t in zip(Rs, targets))
        loss.backward()
        opt.step()
        losses.append(float(loss))
    assert losses[-1] < 0.5 * losses[0], losses[::8]


def test_two_renders_of_one_renderer_in_one_loss(ctx):
    """R(tf1, g) and R(tf2, g) of the same renderer summed into one loss:
    each autograd call keeps its own checkpoints, so the gradients equal those
    of two separate renderers (which cannot share a workspace)."""
    from render3d_amd.autograd import DiffRenderer
    n, W, H = 32, 24, 24
    vol = ctx.synth_sheets(n, n, n, 2, 77)
    v = ctx.volume(n, n, n)
    v.upload(vol)
    view = scene.frame_params(scene.orbit_camera(40.0, 10.0, 1.8), W, H, step_voxels=0.5)
    g0 = torch.ones((8, 8, 8), dtype=torch.float32, device="cuda")
    tf1_0 = torch.from_numpy(gaussian_tf().astype(np.float32)).cuda()
    tf2_0 = torch.full((256, 4), 0.35, dtype=torch.float32, device="cuda")
    w1 = torch.linspace(0.0, 1.0, H * W * 4, device="cuda").reshape(H, W, 4)
    w2 = torch.linspace(1.0, -0.5, H * W * 4, device="cuda").reshape(H, W, 4)

    def grads(renderers):
        tf1 = tf1_0.clone().requires_grad_(True)
        tf2 = tf2_0.clone().requires_grad_(True)
        g = (g0 * 1.1).requires_grad_(True)
        loss = (renderers[0](tf1, g) * w1).sum() + (renderers[1](tf2, g) * w2).sum()
        loss.backward()
        torch.cuda.synchronize()
        return tf1.grad.cpu().numpy(), tf2.grad.cpu().numpy(), g.grad.cpu().numpy()

    R = DiffRenderer(ctx, v, [view], W, H)
    shared = grads([R, R])
    separate = grads([DiffRenderer(ctx, v, [view], W, H), DiffRenderer(ctx, v, [view], W, H)])
    for a, b in zip(shared, separate):
        assert np.abs(b).max() > 0
        assert np.array_equal(a, b) or rel(a, b) <= 1e-5, rel(a, b)  # grid sums use float atomics


def test_full_scale_gradient_in_a_window(ctx):
    """1024^3 volume decoded from its chunks (4 levels), a 128^3 grid and one
    1920x1080 view: the GPU adjoint of a linear loss on a 16x16 ray window
    against the oracle's FP64 adjoint of the same rays and FP64 central
    differences (h = 1e-3) on 32 entries."""
    from render3d_amd.autograd import DiffRenderer
    from test_oracle_adjoint import gaussian_tf

    n, W, H = 1024, 1920, 1080
    step = torch.arange(46, dtype=torch.float32, device="cuda") * 2 + 1
    lin = ctx.synth_sheets(n, n, n, 48, 20261203)
    levels, headers = ctx.encode_chunks(lin, n, n, n, step)
    del lin
    flat = torch.empty((levels.shape[0], 4096), dtype=torch.uint8, device="cuda")
    ctx.decode(ctx.make_batch(headers=headers, step_tables=step, out=flat, levels=levels))
    vol = ctx.volume(n, n, n, 4)
    vol.build_from_chunks(flat)
    del levels, headers, flat
    torch.cuda.synchronize()
    lv = [(d, vol.read_level(l)) for l, d in enumerate(vol.level_dims)]
    ov = orc.OracleVolume(None, n, n, n, levels=lv, occ=(vol.odim, vol.read_occupancy()))
    del lv
    g = 128
    zz, yy, xx = np.meshgrid(*(np.arange(g),) * 3, indexing="ij")
    grid = (1.0 + 0.2 * np.sin(xx / 9.0) * np.cos(yy / 7.0 + zz / 11.0)).astype(np.float32)
    gd = (g, g, g)
    tf = gaussian_tf().astype(np.float32)
    view = scene.frame_params(scene.orbit_camera(45.0, 10.0, 2.0), W, H, step_voxels=1.0)
    win = (952, 532, 968, 548)
    rng = np.random.default_rng(4)
    dimg = np.zeros((H, W, 4), np.float32)
    dimg[win[1]:win[3], win[0]:win[2]] = rng.standard_normal((win[3] - win[1], win[2] - win[0], 4))
    R = DiffRenderer(ctx, vol, [view], W, H, alpha_cutoff=0.98)
    t_tf, t_g = torch.from_numpy(tf).cuda(), torch.from_numpy(grid).cuda()
    img = R.forward_raw(t_tf, t_g)
    a, b = R.adjoint_raw(t_tf, t_g, torch.from_numpy(dimg).cuda())
    torch.cuda.synchronize()
    gtf = a.cpu().numpy().astype(np.float64)
    gg = b.cpu().numpy().reshape(-1).astype(np.float64)
    tfd, gdd = tf.astype(np.float64), grid.astype(np.float64).reshape(-1)
    ref, rtf, rg = orc.raycast_diff(ov, tfd, gdd, gd, [view], W, H, 0.98, dimg=dimg.astype(np.float64), window=win)
    sub = (slice(win[1], win[3]), slice(win[0], win[2]))
    err = np.abs(img.cpu().numpy()[sub] - ref[sub])
    assert np.mean(err < 1e-4) > 0.99 and err.max() < 1e-2, (np.mean(err < 1e-4), err.max())
    assert rel(gtf, rtf) <= 1e-3, rel(gtf, rtf)
    assert rel(gg, rg) <= 1e-3, rel(gg, rg)

    def L():
        im, _, _ = orc.raycast_diff(ov, tfd, gdd, gd, [view], W, H, 0.98, window=win)
        return float(np.sum(im[sub] * dimg[sub]))

    cand_tf = np.argwhere(np.abs(gtf) > 1e-6)
    cand_g = np.nonzero(np.abs(gg) > 1e-6)[0]
    picks = [("tf", tuple(c)) for c in cand_tf[rng.choice(len(cand_tf), 16, replace=False)]]
    picks += [("g", int(i)) for i in rng.choice(cand_g, 16, replace=False)]
    scale = max(np.abs(gtf).max(), np.abs(gg).max())
    checked = 0
    for kind, idx in picks:
        arr = tfd if kind == "tf" else gdd
        base = arr[idx]
        vals = []
        for h in (1e-3, 5e-4):
            arr[idx] = base + h
            lp = L()
            arr[idx] = base - h
            lm = L()
            arr[idx] = base
            vals.append((lp - lm) / (2 * h))
        if abs(vals[0] - vals[1]) > 1e-5 * max(scale, 1.0) + 1e-3 * abs(vals[0]):
            continue  # a gate, the cutoff or a clamp within h of this entry
        got = gtf[idx] if kind == "tf" else gg[idx]
        checked += 1
        assert abs(got - vals[0]) <= 1e-3 * max(abs(vals[0]), abs(got)) + 1e-5 * scale, (kind, idx, got, vals[0])
    print(f"full-scale window: {checked} of 32 entries checked against FP64 central differences")
    assert checked >= 20
