"""Streaming residency from 3ddct shard stores on the GPU through the C ABI:
r3d_hip_bricks_build_from_shards, then r3d_hip_bricks_stream frames whose
fills decode each brick from its level's shard straight into the slot.
Checked frame by frame against the oracle's policy with unknown maxima
(oracle/stream_policy.py, known=..., floor=5) and against the oracle's shard
reader (r3do_shard_decode_region, itself pinned byte for byte to the
reference reader): the page table (entries, maxima, bricks below the noise
floor left unpublished), every slot's voxels, the statistics, and a
bricks-LOD render of the streamed tables against the oracle's raycast_bricks.
The shards are the pyramid of a sheet volume encoded on the GPU (qstep 8)."""
import math

import numpy as np
import pytest

import render3d_amd as r3d
from oracle import oracle as orc
from oracle import stream_policy as sp
from render3d_amd import scene, workloads

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

DIMS = (384, 256, 256)  # x, y, z: 12 + 2 + 1 + 1 bricks over four levels
NLEV, ABPA = 4, 2  # 8 slots for 16 bricks
EMPTY = (2, 1, 1)  # a level-0 brick of zeros: filled once, never published
QSTEP = 8.0


def _shard_of_level(ctx, lev, dims):
    """One shard (at 0, 0, 0) holding every chunk of a level, encoded on the GPU."""
    nx, ny, nz = dims
    step = torch.from_numpy(workloads.step_table_qstep(QSTEP)).cuda()
    vol = torch.from_numpy(np.ascontiguousarray(lev, np.uint8)).cuda()
    levels, hdr = ctx.encode_chunks(vol, nx, ny, nz, step)
    data, offs, lens = ctx.encode_blobs(levels, hdr, QSTEP)
    data, offs, lens = data.cpu().numpy(), offs.cpu().numpy(), lens.cpu().numpy()
    cx, cy = nx // 16, ny // 16
    blobs = {}
    for c in range(offs.size):
        bx, by, bz = c % cx, (c // cx) % cy, c // (cx * cy)
        blobs[(bz * 64 + by) * 64 + bx] = data[offs[c]:offs[c] + lens[c]]
    return orc.shard_pack(blobs)


@pytest.fixture(scope="module")
def zscene(ctx):
    nx, ny, nz = DIMS
    vol = orc.synth_sheets(nx, ny, nz, 6, 91).reshape(nz, ny, nx)
    gx, gy, gz = EMPTY
    vol[gz * 128:(gz + 1) * 128, gy * 128:(gy + 1) * 128, gx * 128:(gx + 1) * 128] = 0
    vol = np.ascontiguousarray(vol).reshape(-1)
    pyr = orc.pyramid(vol, nx, ny, nz, NLEV)
    shards, decoded = [], []
    for dims, lev in pyr:
        sh = _shard_of_level(ctx, lev, dims)
        x, y, z = dims
        rc, dst, counts, _ = orc.shard_decode_region([(sh, 0, 0, 0)], (z, y, x), (0, 0, 0, z, y, x))
        assert rc == 0 and counts[2] == 0
        shards.append(sh)
        decoded.append((dims, dst.reshape(-1)))
    bmax = sp.brick_maxima(decoded)
    dims = [d for d, _ in decoded]
    nent = bmax.size
    e_empty = (gz * 2 + gy) * 3 + gx
    assert bmax[e_empty] < sp.NOISE_FLOOR and (bmax[:e_empty] >= sp.NOISE_FLOOR).all()
    resident = np.zeros(nent, np.uint8)
    resident[-1] = 1
    dshards = [torch.from_numpy(s).cuda() for s in shards]
    return dict(dshards=dshards, decoded=decoded, dims=dims, bmax=bmax, resident=resident, nent=nent,
                e_empty=e_empty)


def _brick(z, e, pol):
    """The decoded voxels of brick e, zero-padded to 128^3."""
    l, bx, by, bz = pol.locate(e)
    (x, y, zz), arr = z["decoded"][l]
    a = arr.reshape(zz, y, x)[bz * 128:(bz + 1) * 128, by * 128:(by + 1) * 128, bx * 128:(bx + 1) * 128]
    out = np.zeros((128, 128, 128), np.uint8)
    out[:a.shape[0], :a.shape[1], :a.shape[2]] = a
    return out


def _fresh(ctx, z):
    b = r3d.Bricks(ctx, *DIMS, NLEV, ABPA)
    b.build_from_shards([[(s, 0, 0, 0)] for s in z["dshards"]], z["resident"])
    pol = sp.StreamPolicy(z["dims"], ABPA ** 3, z["bmax"], z["resident"], known=z["resident"], floor=sp.NOISE_FLOOR)
    return b, pol


def _header(z):
    ov_page = orc.brick_page_words(z["dims"])
    ob = orc.bricks_build(orc.OracleVolume(None, *DIMS, levels=[(d, a) for d, a in z["decoded"]],
                                           occ=((1, 1, 1), np.zeros(1, np.uint8))), 1, np.zeros(z["nent"], np.uint8))
    assert ob.page.size == ov_page
    return ob.page[:64]


def _compare(b, pol, z, hdr, tag):
    torch.cuda.synchronize()
    page = b.read_page()
    want = np.concatenate([hdr, pol.entries()])
    bad = np.nonzero(page != want)[0]
    assert bad.size == 0, f"{tag}: page words {bad[:8]} {page[bad[:8]]} != {want[bad[:8]]}"
    slots = b.read_slots(0, ABPA ** 3)
    for slot, e in enumerate(pol.slot_brick):
        if e != sp.INV:
            assert np.array_equal(slots[slot], _brick(z, e, pol)), f"{tag}: slot {slot} (entry {e})"
    st = b.stream_stats()
    assert st["hot"] == sum(e != sp.INV for e in pol.slot_brick)
    assert st["lod_wanted"] == pol.lod_wanted and st["lod_requests"] == pol.lod_requests, tag
    return st


def _path(i, n):
    a = 2.0 * math.pi * i / n
    eye = [0.5 + 0.32 * math.cos(a), 0.33 + 0.05 * math.sin(3 * a), 0.33 + 0.32 * math.sin(a)]
    fwd = np.array([-math.sin(a), -0.15, math.cos(a)])
    fwd /= np.linalg.norm(fwd)
    return [float(v) for v in np.float32(eye)], [float(v) for v in np.float32(fwd)]


def test_shard_stream_frames_match_policy(ctx, zscene):
    z = zscene
    b, pol = _fresh(ctx, z)
    hdr = _header(z)
    _compare(b, pol, z, hdr, "built")
    ht = float(np.float32(math.tan(math.radians(30.0))))
    fills = evicts = 0
    seen_empty = False
    for i in range(24):
        eye, fwd = _path(i, 24)
        h = (96, 60, 200)[i % 3]
        pc = float(np.float32(2.0 * ht / h))
        gate = (0.0, 0.2)[i % 2]
        budget = (2, 6, 1, 32, 40, 0)[i % 6]
        b.stream(eye, fwd, ht, pc, gate, budget)
        filled, evicted = pol.stream(eye, fwd, ht, pc, gate, budget)
        st = _compare(b, pol, z, hdr, f"frame {i}")
        assert (st["inflight"], st["evicted_last"]) == (len(filled), len(evicted)), i
        fills += len(filled)
        evicts += len(evicted)
        seen_empty |= any(e == z["e_empty"] for e, _ in filled)
    print(f"filled {fills} evicted {evicts}")
    assert fills >= 20 and evicts >= 8
    assert seen_empty and pol.known[z["e_empty"]]  # filled once, measured, not published
    b.close()


def test_shard_streamed_tables_render_as_oracle(ctx, zscene):
    z = zscene
    b, pol = _fresh(ctx, z)
    hdr = _header(z)
    ht = float(np.float32(math.tan(math.radians(30.0))))
    for i in range(6):
        eye, fwd = _path(i, 12)
        b.stream(eye, fwd, ht, float(np.float32(2.0 * ht / 96)), 0.0, 6)
        pol.stream(eye, fwd, ht, np.float32(2.0 * ht / 96), 0.0, 6)
        torch.cuda.synchronize()  # the fills' maxima are applied at the next begin
    _compare(b, pol, z, hdr, "rendered")
    atlas = np.zeros(ABPA ** 3 * orc.SLOT_VOX, np.uint8)
    for slot, e in enumerate(pol.slot_brick):
        if e != sp.INV:
            atlas[slot * orc.SLOT_VOX:(slot + 1) * orc.SLOT_VOX] = _brick(z, e, pol).reshape(-1)
    ob = orc.OracleBricks(np.concatenate([hdr, pol.entries()]), atlas, ABPA)
    lut = scene.tf_ramp_40_200()
    W, H = 112, 80
    views = [scene.frame_params(scene.orbit_camera(40.0, 15.0, 1.2), W, H, step_voxels=0.5,
                                gate=scene.skip_gate(lut), brick_mode=b.brick_mode),
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


def test_shard_source_rejections(ctx, zscene):
    b = r3d.Bricks(ctx, *DIMS, NLEV, ABPA)
    with pytest.raises(ValueError):
        b.build_from_shards([[]] * (NLEV - 1))
    with pytest.raises(r3d.R3DError):  # more resident bricks than slots
        b.build_from_shards([[]] * NLEV)
    # no shards at all: every brick decodes to zeros, none is published
    res = np.zeros(zscene["nent"], np.uint8)
    res[-1] = 1
    b.build_from_shards([[]] * NLEV, res)
    torch.cuda.synchronize()
    page = b.read_page()
    assert (page[64:] == 0xFFFFFFFF).all() and not b.read_slot(0).any()
    b.close()


def test_build_from_shards_levels_off_the_chunk_grid(ctx):
    """Levels whose dims are not multiples of 16 (their fills take one region
    decode per brick; the others one batched decode per level): every brick
    of a 200 x 136 x 120 three-level pyramid, made resident at build, equals
    the oracle reader's decode of its box, zero past the level."""
    dims = (200, 136, 120)
    nx, ny, nz = dims
    vol = orc.synth_sheets(nx, ny, nz, 4, 13)
    pyr = orc.pyramid(vol, nx, ny, nz, 3)
    shards, decoded = [], []
    for (x, y, z), lev in pyr:
        px, py, pz = -(-x // 16) * 16, -(-y // 16) * 16, -(-z // 16) * 16  # encoded padded to whole chunks
        pad = np.zeros((pz, py, px), np.uint8)
        pad[:z, :y, :x] = lev.reshape(z, y, x)
        sh = _shard_of_level(ctx, pad.reshape(-1), (px, py, pz))
        rc, dst, counts, _ = orc.shard_decode_region([(sh, 0, 0, 0)], (z, y, x), (0, 0, 0, z, y, x))
        assert rc == 0
        shards.append(torch.from_numpy(sh).cuda())
        decoded.append(((x, y, z), dst.reshape(-1)))
    ldims = [d for d, _ in decoded]
    pol = sp.StreamPolicy(ldims, 27, sp.brick_maxima(decoded))
    b = r3d.Bricks(ctx, nx, ny, nz, 3, 3)
    b.build_from_shards([[(s, 0, 0, 0)] for s in shards])
    torch.cuda.synchronize()
    z = {"decoded": decoded}
    slots = b.read_slots(0, pol.nent)
    page = b.read_page()
    for e in range(pol.nent):
        assert np.array_equal(slots[e], _brick(z, e, pol)), e
        m = int(_brick(z, e, pol).max())
        assert page[64 + e] == (e | m << 24 if m >= sp.NOISE_FLOOR else 0xFFFFFFFF), e
    b.close()
