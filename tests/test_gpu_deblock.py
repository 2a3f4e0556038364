"""Deblocking on the GPU through the C ABI: the volume and chunk forms
against the reference tools/3ddct/deblock.c's outputs
(tests/golden/dct3d_deblock.npz) bit for bit, and against the oracle on a
larger decoded volume."""
import os

import numpy as np
import pytest

from oracle import oracle as orc
from render3d_amd import workloads

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

D = np.load(os.path.join(os.path.dirname(__file__), "golden", "dct3d_deblock.npz"))


def vol(name):
    if name == "edge":
        return np.where(np.arange(64)[None, None, :] < 32, 40, 200).astype(np.uint8) * np.ones((64, 64, 1), np.uint8)
    return D[name]


def test_volume_form_matches_the_reference(ctx):
    for k, name in enumerate(D["case_vol"]):
        v = torch.from_numpy(vol(str(name)).copy()).cuda()
        ctx.deblock(v, float(D["case_step"][k]), float(D["case_strength"][k]))
        torch.cuda.synchronize()
        want = D["case_out"][D["case_off"][k]:D["case_off"][k + 1]].reshape(v.shape)
        assert np.array_equal(v.cpu().numpy(), want), (name, D["case_step"][k], D["case_strength"][k])


def test_chunk_form_matches_the_reference(ctx):
    a = D["a"]
    chunks = a.reshape(4, 16, 4, 16, 4, 16).transpose(0, 2, 4, 1, 3, 5).reshape(64, 4096).copy()
    for k in range(D["ck_center"].size):
        c, m = int(D["ck_center"][k]), int(D["ck_mask"][k])
        cz, cy, cx = c // 16, (c // 4) % 4, c % 4
        nb = [(cz, cy, cx - 1), (cz, cy, cx + 1), (cz, cy - 1, cx), (cz, cy + 1, cx), (cz - 1, cy, cx), (cz + 1, cy, cx)]
        # batch: the centre first, then its neighbours
        rows = [chunks[c]] + [chunks[(z * 4 + y) * 4 + x] for f, (z, y, x) in enumerate(nb) if (m >> f) & 1]
        idx = -np.ones((len(rows), 6), np.int32)
        j = 1
        for f in range(6):
            if (m >> f) & 1:
                idx[0, f] = j
                j += 1
        out = ctx.deblock_chunks(torch.from_numpy(np.stack(rows)).cuda(), torch.from_numpy(idx).cuda(),
                                 float(D["ck_step"][k]), float(D["ck_strength"][k]))
        torch.cuda.synchronize()
        assert np.array_equal(out[0].cpu().numpy(), D["ck_out"][k]), k
        assert torch.equal(out[1:].cpu(), torch.from_numpy(np.stack(rows[1:]))), k  # no neighbours: unchanged


def test_decoded_volume_against_the_oracle(ctx):
    """A 192 x 160 x 208 region of the q=32 sheet volume decode (seams in
    every direction, dims not multiples of 16 past the first seam), in both
    threshold modes."""
    n = 256
    step = torch.from_numpy(workloads.step_table_qstep(32.0)).cuda()
    v = ctx.synth_sheets(n, n, n, 20, 5)
    levels, headers = ctx.encode_chunks(v, n, n, n, step)
    flat = ctx.decode_chunks(headers=headers, step_tables=step, levels=levels)
    full = flat.view(16, 16, 16, 16, 16, 16).permute(0, 3, 1, 4, 2, 5).reshape(n, n, n)
    base = full[7:199, 30:190, 11:219].contiguous()
    for s, t in ((32.0, 1.0), (0.0, 1.3)):
        g = base.clone()
        ctx.deblock(g, s, t)
        torch.cuda.synchronize()
        want = orc.deblock(base.cpu().numpy(), s, t)
        assert (want != base.cpu().numpy()).sum() > 1000
        assert np.array_equal(g.cpu().numpy(), want), (s, t)


def test_no_op_cases(ctx):
    v = torch.from_numpy(D["a"].copy()).cuda()
    ctx.deblock(v, 24.0, 0.0)  # strength 0
    small = torch.from_numpy(D["a"][:17, :17, :17].copy()).cuda()  # no interior plane with a layer past it
    ctx.deblock(small, 24.0, 1.0)
    torch.cuda.synchronize()
    assert torch.equal(v.cpu(), torch.from_numpy(D["a"]))
    assert torch.equal(small.cpu(), torch.from_numpy(D["a"][:17, :17, :17].copy()))


def test_every_level_pair_across_a_seam(ctx):
    """Every pair (p0, q0) of u8 values across an x seam, with flat sides
    (p1 = p0, q1 = q0), at step 127.5 (alpha 255, beta 153), so every step
    |q0 - p0| <= 254 is a seam: all 511 entries of the tile kernel's nudge
    table and the byte rounding of every p0 + d0, plus the y and z seams the
    pattern makes, against the oracle. The same bytes at an odd address take
    the per-pass kernels."""
    z, y = np.meshgrid(np.arange(256), np.arange(256), indexing="ij")
    base = np.empty((256, 256, 32), np.uint8)
    base[:, :, :16] = y[:, :, None]
    base[:, :, 16:] = z[:, :, None]
    want = orc.deblock(base, 127.5, 1.0)
    seam = want[:, :, 14:18].astype(np.int32)
    assert (seam[:, :, 1] != y).sum() > 60000  # nearly every pair was nudged
    g = torch.from_numpy(base.copy()).cuda()
    assert g.data_ptr() % 16 == 0
    ctx.deblock(g, 127.5, 1.0)
    buf = torch.zeros(base.size + 16, dtype=torch.uint8, device="cuda")
    odd = buf[1:1 + base.size].view(base.shape)
    odd.copy_(torch.from_numpy(base))
    ctx.deblock(odd, 127.5, 1.0)
    torch.cuda.synchronize()
    assert np.array_equal(g.cpu().numpy(), want)
    assert np.array_equal(odd.cpu().numpy(), want)


def test_both_kernel_forms_against_the_oracle(ctx):
    """A row length that is a multiple of 128 (the tile pass's last tile holds
    only the two voxels before the end of each row) on a 16-byte aligned
    volume, and the same bytes at an odd address, which takes the per-pass
    kernels."""
    n = 256
    step = torch.from_numpy(workloads.step_table_qstep(32.0)).cuda()
    v = ctx.synth_sheets(n, n, n, 20, 9)
    levels, headers = ctx.encode_chunks(v, n, n, n, step)
    flat = ctx.decode_chunks(headers=headers, step_tables=step, levels=levels)
    full = flat.view(16, 16, 16, 16, 16, 16).permute(0, 3, 1, 4, 2, 5).reshape(n, n, n)
    base = full[3:70, 14:99, :].contiguous()
    want = orc.deblock(base.cpu().numpy(), 32.0, 1.0)
    assert (want != base.cpu().numpy()).sum() > 1000
    g = base.clone()
    assert g.data_ptr() % 16 == 0
    ctx.deblock(g, 32.0, 1.0)
    buf = torch.zeros(base.numel() + 16, dtype=torch.uint8, device="cuda")
    odd = buf[1:1 + base.numel()].view(base.shape)
    odd.copy_(base)
    ctx.deblock(odd, 32.0, 1.0)
    torch.cuda.synchronize()
    assert np.array_equal(g.cpu().numpy(), want)
    assert np.array_equal(odd.cpu().numpy(), want)
    assert int(buf[0]) == 0 and int(buf[-15:].sum()) == 0
