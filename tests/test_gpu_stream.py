"""Streaming residency on the GPU through the C ABI: r3d_hip_bricks_stream
(the view-cone collect), _begin/_box/_point/_submit and _stats against the
oracle's restatement of the reference's pump (oracle/stream_policy.py),
frame by frame over a camera path on an atlas smaller than the pyramid:
the whole page table (entries and their maxima), every slot's voxels
(against the oracle pyramid's bricks), the statistics, and a bricks-LOD
render of the streamed tables against the oracle's raycast_bricks."""
import math

import numpy as np
import pytest

import render3d_amd as r3d
from oracle import oracle as orc
from oracle import stream_policy as sp
from render3d_amd import scene

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

DIMS = (384, 256, 256)  # x, y, z: 12 + 2 + 1 + 1 bricks over four levels
NLEV, ABPA = 4, 2  # 8 slots for 16 bricks
GATED = (2, 1, 1)  # a level-0 brick kept below the noise floor


@pytest.fixture(scope="module")
def sscene(ctx):
    nx, ny, nz = DIMS
    vol = orc.synth_sheets(nx, ny, nz, 6, 77).reshape(nz, ny, nx)
    gx, gy, gz = GATED
    blk = vol[gz * 128:(gz + 1) * 128, gy * 128:(gy + 1) * 128, gx * 128:(gx + 1) * 128]
    np.minimum(blk, 3, out=blk)
    vol = np.ascontiguousarray(vol).reshape(-1)
    ov = orc.OracleVolume(vol, nx, ny, nz, nlev=NLEV)
    dims = [d for d, _ in ov.levels]
    full = orc.bricks_build(ov, 3)  # every brick resident: slot e holds entry e
    nent = full.page.size - 64
    bmax = (full.page[64:] >> 24).astype(np.int64)
    assert np.array_equal(bmax, sp.brick_maxima(ov.levels))
    assert bmax[(gz * 2 + gy) * 3 + gx] <= 3
    resident = np.zeros(nent, np.uint8)
    resident[-1] = 1
    dvol = torch.from_numpy(vol).cuda()
    return dict(vol=dvol, ov=ov, dims=dims, full=full, bmax=bmax, resident=resident, nent=nent)


def _fresh(ctx, s):
    b = r3d.Bricks(ctx, *DIMS, NLEV, ABPA)
    b.build(s["vol"], s["resident"])
    pol = sp.StreamPolicy(s["dims"], ABPA ** 3, s["bmax"], s["resident"])
    return b, pol


def _compare(b, pol, s, tag):
    torch.cuda.synchronize()
    page = b.read_page()
    want = np.concatenate([s["full"].page[:64], pol.entries()])
    bad = np.nonzero(page != want)[0]
    assert bad.size == 0, f"{tag}: page words {bad[:8]} {page[bad[:8]]} != {want[bad[:8]]}"
    slots = b.read_slots(0, ABPA ** 3)
    for slot, e in enumerate(pol.slot_brick):
        if e == sp.INV:
            continue
        ref = s["full"].atlas[e * orc.SLOT_VOX:(e + 1) * orc.SLOT_VOX].reshape(128, 128, 128)
        assert np.array_equal(slots[slot], ref), f"{tag}: slot {slot} (entry {e})"
    st = b.stream_stats()
    assert st["nb"] == pol.nent and st["hot_cap"] == pol.nslots and st["nlevels"] == NLEV
    assert st["hot"] == sum(e != sp.INV for e in pol.slot_brick)
    assert st["lod_wanted"] == pol.lod_wanted and st["lod_requests"] == pol.lod_requests, tag
    return st


def _path(i, n):
    a = 2.0 * math.pi * i / n
    eye = [0.5 + 0.32 * math.cos(a), 0.33 + 0.05 * math.sin(3 * a), 0.33 + 0.32 * math.sin(a)]
    fwd = np.array([-math.sin(a), -0.15, math.cos(a)])
    fwd /= np.linalg.norm(fwd)
    return [float(v) for v in np.float32(eye)], [float(v) for v in np.float32(fwd)]


def test_stream_cone_frames_match_policy(ctx, sscene):
    b, pol = _fresh(ctx, sscene)
    _compare(b, pol, sscene, "built")
    ht = float(np.float32(math.tan(math.radians(30.0))))
    fills = evicts = 0
    for i in range(24):
        eye, fwd = _path(i, 24)
        h = (96, 60, 200)[i % 3]
        pc = float(np.float32(2.0 * ht / h))
        gate = (0.0, 0.2)[i % 2]
        budget = (2, 6, 1, 32, 40, 0)[i % 6]
        sz0, sd = (0, 0) if i % 4 else (100, 90)
        b.stream(eye, fwd, ht, pc, gate, budget, slice_z0=sz0, slice_depth=sd)
        filled, evicted = pol.stream(eye, fwd, ht, pc, gate, budget, slice_z0=sz0, slice_depth=sd)
        st = _compare(b, pol, sscene, f"frame {i}")
        assert (st["inflight"], st["evicted_last"]) == (len(filled), len(evicted)), i
        fills += len(filled)
        evicts += len(evicted)
        assert st["decoded"] == fills and st["evicted"] == evicts
    print(f"filled {fills} evicted {evicts}")
    assert fills >= 20 and evicts >= 8
    gated = pol.entry(0, *GATED)
    assert pol.brick_slot[gated] == sp.INV
    b.close()


def test_stream_box_and_point_collects(ctx, sscene):
    b, pol = _fresh(ctx, sscene)
    frames = [
        ("box", (0.0, 0.0, 0.0), (1.0, 0.66, 0.66), 2.0 / 384, 0.0, 8),
        ("box", (0.1, 0.05, 0.2), (0.5, 0.3, 0.6), 1.0 / 384, 0.0, 3),
        ("point", (0.9, 0.6, 0.1), 0, 0.0, 4),
        ("both", (0.4, 0.2, 0.1), (0.95, 0.6, 0.3), 0.5 / 384, 0.1, 32),
        ("point", (0.01, 0.01, 0.01), 5, 0.0, 2),  # a level past the pyramid clamps to the coarsest
    ]
    for k, f in enumerate(frames):
        assert b.stream_begin() and pol.begin()
        if f[0] in ("box", "both"):
            b.stream_box(f[1], f[2], f[3], f[4])
            pol.box(f[1], f[2], f[3], f[4])
        if f[0] == "point":
            b.stream_point(f[1], f[2], f[3])
            pol.point(f[1], f[2], f[3])
        if f[0] == "both":
            for p in ((0.7, 0.5, 0.6), (0.2, 0.6, 0.5)):
                b.stream_point(p, 0, f[4])
                pol.point(p, 0, f[4])
        b.stream_submit(f[-1])
        filled, evicted = pol.submit(f[-1])
        st = _compare(b, pol, sscene, f"collect {k}")
        assert (st["inflight"], st["evicted_last"]) == (len(filled), len(evicted))
    # collects and a submit without an open pass change nothing
    page, st = b.read_page(), b.stream_stats()
    b.stream_box((0.0, 0.0, 0.0), (1.0, 1.0, 1.0), 0.0, 0.0)
    b.stream_submit(8)
    torch.cuda.synchronize()
    assert np.array_equal(b.read_page(), page) and b.stream_stats() == st
    b.close()


def test_streamed_tables_render_as_oracle(ctx, sscene):
    b, pol = _fresh(ctx, sscene)
    ht = float(np.float32(math.tan(math.radians(30.0))))
    for i in range(6):
        eye, fwd = _path(i, 12)
        b.stream(eye, fwd, ht, float(np.float32(2.0 * ht / 96)), 0.0, 6)
        pol.stream(eye, fwd, ht, np.float32(2.0 * ht / 96), 0.0, 6)
    _compare(b, pol, sscene, "rendered")
    atlas = np.zeros(ABPA ** 3 * orc.SLOT_VOX, np.uint8)
    full = sscene["full"]
    for slot, e in enumerate(pol.slot_brick):
        if e != sp.INV:
            atlas[slot * orc.SLOT_VOX:(slot + 1) * orc.SLOT_VOX] = full.atlas[e * orc.SLOT_VOX:(e + 1) * orc.SLOT_VOX]
    ob = orc.OracleBricks(np.concatenate([full.page[:64], pol.entries()]), atlas, ABPA)
    lut = scene.tf_ramp_40_200()
    gate = scene.skip_gate(lut)
    W, H = 112, 80
    views = [scene.frame_params(scene.orbit_camera(40.0, 15.0, 1.2), W, H, step_voxels=0.5, gate=gate,
                                brick_mode=b.brick_mode),
             scene.frame_params(scene.orbit_camera(200.0, -10.0, 0.9), W, H, mode=r3d.MODE_MIP,
                                brick_mode=b.brick_mode)]
    ctx.set_transfer(lut)
    for v in views:
        out = torch.zeros((H, W, 4), dtype=torch.float32, device="cuda")
        res = b.raycast([v], W, H, alpha_cutoff=0.99, out=out, stats=True)
        ref, cnt, _ = orc.raycast_bricks(ob, scene.tf_unorm(lut), [v], W, H, alpha_cutoff=0.99)
        err = float(np.max(np.abs(res["rgba"].cpu().numpy() - ref)))
        assert err <= 1e-5, err
        st = res["stats"]
        assert (st["samples"], st["skips"], st["shaded"]) == (cnt["samples"], cnt["skips"], cnt["shaded"])
        assert cnt["samples"] > 1000
    b.close()


def test_stream_before_build_is_refused(ctx):
    b = r3d.Bricks(ctx, 256, 128, 128, 2, 1)
    assert not b.stream_begin()
    with pytest.raises(r3d.R3DError):
        b.stream([0.5, 0.25, 0.25], [0.0, 0.0, 1.0], 0.5, 0.01, 0.0, 4)
    b.close()
