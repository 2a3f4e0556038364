"""Batched path through the C ABI: raycast plans (every launch shape gives the
immediate call's image and counters bit for bit), the autotuner, and one HIP
graph holding a sparse-batch decode and a multi-view render, replayed with
eager work in between."""
import numpy as np
import pytest

from oracle import oracle as orc
import render3d_amd as r3d
from render3d_amd import scene, workloads

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

SHAPES = [(8, 8), (16, 4), (16, 8), (16, 16)]


@pytest.fixture(scope="module")
def small_scene(ctx):
    n = 96
    vol = ctx.volume(n, n, n, 3)
    vol.upload(ctx.synth_sheets(n, n, n, 4, 99))
    vol.build()
    lut = scene.tf_ramp_40_200()
    ctx.set_transfer(lut)
    gate = scene.skip_gate(lut)
    sizes = ((2, 40, 24), (2, 64, 64), (1, 100, 70))
    views, w, h = workloads.mixed_views(sizes, atlas_width=160, step_voxels=0.75, gate=gate)
    views[1].mode = r3d.MODE_MIP
    views[3].mode = r3d.MODE_FLAT
    return vol, views, w, h


def immediate(ctx, vol, views, w, h):
    counters = torch.zeros(4, dtype=torch.int64, device="cuda")
    out = torch.zeros((h, w, 4), dtype=torch.float32, device="cuda")  # pixels no view covers stay 0
    r = ctx.raycast(vol, views, w, h, rgba8=True, counters=counters, out=out)
    torch.cuda.synchronize()
    return r["rgba"].cpu().numpy(), r["rgba8"].cpu().numpy(), counters.cpu().numpy()


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("bpc", [0, 1, 4])
def test_plan_matches_immediate_call(ctx, small_scene, shape, bpc):
    vol, views, w, h = small_scene
    ref, ref8, cref = immediate(ctx, vol, views, w, h)
    counters = torch.zeros(4, dtype=torch.int64, device="cuda")
    plan = ctx.raycast_plan(vol, views, w, h, rgba8=True, counters=counters)
    plan.set_tuning(*shape, bpc)
    plan.launch()
    plan.launch()
    torch.cuda.synchronize()
    assert np.array_equal(plan.rgba.cpu().numpy().view(np.uint32), ref.view(np.uint32))
    assert np.array_equal(plan.rgba8.cpu().numpy(), ref8)
    assert np.array_equal(counters.cpu().numpy(), 2 * cref)  # counters accumulate per launch
    plan.close()


def test_immediate_call_honours_tuning(ctx, small_scene):
    vol, views, w, h = small_scene
    ref, _, cref = immediate(ctx, vol, views, w, h)
    try:
        ctx.set_raycast_tuning(16, 8, 2)
        got, _, cgot = immediate(ctx, vol, views, w, h)
    finally:
        ctx.set_raycast_tuning(8, 8, 0)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
    assert np.array_equal(cgot, cref)


def test_tuning_is_validated(ctx, small_scene):
    vol, views, w, h = small_scene
    plan = ctx.raycast_plan(vol, views, w, h)
    with pytest.raises(r3d.R3DError):
        plan.set_tuning(8, 16)
    with pytest.raises(r3d.R3DError):
        plan.set_tuning(8, 8, 33)
    with pytest.raises(r3d.R3DError):
        ctx.set_raycast_tuning(32, 32)
    plan.close()


def test_autotuner_keeps_the_fastest_shape(ctx, small_scene):
    vol, views, w, h = small_scene
    plan = ctx.raycast_plan(vol, views, w, h)
    res = plan.tune()
    assert len(res) == 20
    assert {r["tile"] for r in res} == set(SHAPES)
    assert all(r["ms"] > 0 for r in res)
    best = min(res, key=lambda r: r["ms"])
    t = plan.tuning
    assert t["tile"] == best["tile"] and t["blocks_per_cu"] == best["blocks_per_cu"]
    plan.close()


def test_graph_of_sparse_decode_and_plan_replays(ctx, small_scene):
    vol, views, w, h = small_scene
    n = 1536
    offs, pairs, hdr = workloads.sparse_laplacian_batch(n, 5)
    step = workloads.step_table_linear()
    ref_u8, status = orc.decode_batch(workloads.densify(offs, pairs), hdr, step)
    assert status.all()
    d_offs, d_pairs = torch.from_numpy(offs).cuda(), torch.from_numpy(pairs).cuda()
    d_hdr, d_step = torch.from_numpy(hdr).cuda(), torch.from_numpy(step).cuda()
    out = torch.zeros((n, 4096), dtype=torch.uint8, device="cuda")
    batch = ctx.make_batch(headers=d_hdr, step_tables=d_step, out=out, pairs=d_pairs, pair_offsets=d_offs)
    counters = torch.zeros(4, dtype=torch.int64, device="cuda")
    plan = ctx.raycast_plan(vol, views, w, h, rgba8=True, counters=counters)
    plan.set_tuning(16, 8, 4)
    # eager run: reference results
    ctx.decode(batch)
    plan.launch()
    torch.cuda.synchronize()
    assert np.array_equal(out.cpu().numpy(), ref_u8)
    img_ref = plan.rgba.cpu().numpy().copy()
    c_one = counters.cpu().numpy().copy()
    # capture both into one graph on a side stream
    g = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        with torch.cuda.graph(g, stream=s):
            ctx.decode(batch)
            plan.launch()
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    counters.zero_()
    for rep in range(3):
        out.zero_()
        plan.rgba.zero_()
        junk = torch.randn(1 << 20, device="cuda").sum()  # eager work between replays
        ctx.raycast(vol, views[:2], w, h)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert np.array_equal(out.cpu().numpy(), ref_u8), rep
        assert np.array_equal(plan.rgba.cpu().numpy().view(np.uint32), img_ref.view(np.uint32)), rep
        assert np.array_equal(counters.cpu().numpy(), (rep + 1) * c_one), rep
        del junk
    plan.close()


@pytest.mark.parametrize("bpc", [0, 16])
def test_counters_reduce_over_many_workgroups(ctx, small_scene, bpc):
    vol = small_scene[0]
    W, H = 640, 480  # 4,800 8x8 tiles: the counter reduction runs on several workgroups
    view = scene.frame_params(scene.orbit_camera(30.0, 10.0, 1.8), W, H, step_voxels=0.5)
    try:
        ctx.set_raycast_tuning(8, 8, bpc)
        res = ctx.raycast(vol, [view], W, H, px_steps=True, stats=True)
    finally:
        ctx.set_raycast_tuning(8, 8, 0)
    st = res["stats"]
    steps = res["px_steps"].cpu().numpy().astype(np.int64)
    assert st["rays"] == W * H
    assert st["samples"] == int(steps.sum())
