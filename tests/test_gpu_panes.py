"""The pane cache (R3D_RAYCAST_PANE_CACHE, after vkbackend.c:7066-7079,
7144-7159) and the GPU frame time in the stats (r3d_frame_stats.gpu_raycast_ns,
render_types.h:107-119), through the C ABI: unchanged views are not drawn
and keep their pixels, a changed view redraws alone, every scene change
(volume build, transfer function, quality, tuning, invalidate) and layout
change redraws all, a render without the flag resets the cache, and the
image after any mix of cached and redrawn views equals a full render of the
same views against the oracle."""
import numpy as np
import pytest

import render3d_amd as r3d
from oracle import oracle as orc
from render3d_amd import scene

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

N = 64
W, H = 96, 72


@pytest.fixture(scope="module")
def pscene(ctx):
    vol = orc.synth_sheets(N, N, N, 4, 31)
    v = ctx.volume(N, N, N)
    v.upload(vol)
    ov = orc.OracleVolume(vol, N, N, N, nlev=v.nlevels)
    return v, ov, vol


def _views(yaws, frame_index=0, mode=r3d.MODE_FULL):
    w2, h2 = W // 2, H // 2
    orgs = [(0, 0), (w2, 0), (0, h2), (w2, h2)]
    return [scene.frame_params(scene.orbit_camera(y, 10.0, 2.0), w2, h2, view_org=o, frame_index=frame_index,
                               step_voxels=0.75, mode=mode)
            for y, o in zip(yaws, orgs)]


def _draw(ctx, v, views, out, **kw):
    res = ctx.raycast(v, views, W, H, out=out, stats=True, pane_cache=True, **kw)
    torch.cuda.synchronize()
    return res["stats"]


def _full(ctx, v, views):
    out = torch.zeros((H, W, 4), dtype=torch.float32, device="cuda")
    st = ctx.raycast(v, views, W, H, out=out, stats=True)["stats"]
    return out.cpu().numpy(), st


def test_pane_cache_draws_changed_views_only(ctx, pscene):
    v, ov, _ = pscene
    lut = scene.tf_ramp_40_200()
    ctx.set_transfer(lut)
    out = torch.zeros((H, W, 4), dtype=torch.float32, device="cuda")
    yaws = [10.0, 80.0, 150.0, 220.0]
    st = _draw(ctx, v, _views(yaws), out)
    assert st["panes_drawn"] == 4 and st["samples"] > 0 and st["gpu_raycast_ns"] > 0
    first = out.cpu().numpy()
    ref, _, _ = orc.raycast(ov, scene.tf_unorm(lut), _views(yaws), W, H)
    assert float(np.max(np.abs(first - ref))) <= 1e-5
    # nothing changed (frame_index aside): nothing drawn, pixels kept
    out_before = out.clone()
    st = _draw(ctx, v, _views(yaws, frame_index=7), out)
    assert st["panes_drawn"] == 0 and st["samples"] == 0 and st["gpu_raycast_ns"] == 0
    assert torch.equal(out, out_before)
    # one view moves: it alone is drawn; the image equals a full render
    yaws[2] = 170.0
    st = _draw(ctx, v, _views(yaws), out)
    assert st["panes_drawn"] == 1
    full, fst = _full(ctx, v, _views(yaws))
    assert np.array_equal(out.cpu().numpy(), full)
    single, sst = _full(ctx, v, [_views(yaws)[2]])
    assert (st["samples"], st["shaded"], st["rays"]) == (sst["samples"], sst["shaded"], sst["rays"])
    assert fst["panes_drawn"] == 4


def test_pane_cache_scene_changes_redraw_all(ctx, pscene):
    v, _, vol = pscene
    lut = scene.tf_ramp_40_200()
    ctx.set_transfer(lut)
    out = torch.zeros((H, W, 4), dtype=torch.float32, device="cuda")
    views = _views([30.0, 100.0, 190.0, 280.0])
    assert _draw(ctx, v, views, out)["panes_drawn"] == 4
    assert _draw(ctx, v, views, out)["panes_drawn"] == 0
    changes = [
        ("transfer", lambda: ctx.set_transfer(lut)),
        ("quality", lambda: ctx.set_quality(r3d.QUALITY_FULL)),
        ("tuning", lambda: ctx.set_raycast_tuning(8, 8)),
        ("rebuild", lambda: v.upload(vol)),
        ("invalidate", ctx.invalidate_panes),
        ("uncached render", lambda: ctx.raycast(v, views[:1], W, H, out=out)),
    ]
    for name, fn in changes:
        fn()
        assert _draw(ctx, v, views, out)["panes_drawn"] == 4, name
        assert _draw(ctx, v, views, out)["panes_drawn"] == 0, name
    # a layout change (one viewport) redraws every view
    views2 = _views([30.0, 100.0, 190.0, 280.0])
    views2[3].viewport[0] -= 4
    assert _draw(ctx, v, views2, out)["panes_drawn"] == 4
    # another output or other options: drawn again
    out2 = torch.zeros_like(out)
    assert _draw(ctx, v, views2, out2)["panes_drawn"] == 4
    assert _draw(ctx, v, views2, out2, alpha_cutoff=0.9)["panes_drawn"] == 4
    assert _draw(ctx, v, views2, out2, alpha_cutoff=0.9)["panes_drawn"] == 0
    # a mode change is a parameter change of that view
    views3 = list(views2)
    views3[1] = _views([30.0, 100.0, 190.0, 280.0], mode=r3d.MODE_MIP)[1]
    assert _draw(ctx, v, views3, out2, alpha_cutoff=0.9)["panes_drawn"] == 1
    full = torch.zeros_like(out)
    ctx.raycast(v, views3, W, H, out=full, alpha_cutoff=0.9)
    torch.cuda.synchronize()
    assert torch.equal(out2, full)


def test_gpu_time_in_stats(ctx, pscene):
    v, _, _ = pscene
    views = _views([10.0, 80.0, 150.0, 220.0])
    _, st = _full(ctx, v, views)
    assert st["panes_drawn"] == 4
    assert 0 < st["gpu_raycast_ns"] < 1e9
