"""The oracle's deblocking filter (oracle/oracle_deblock.c) against the
reference tools/3ddct/deblock.c: the volume and chunk APIs on the reference's
outputs (tests/golden/dct3d_deblock.npz), and the checks of the reference's
own tools/3ddct/test_deblock.c."""
import os

import numpy as np

from oracle import oracle as orc

D = np.load(os.path.join(os.path.dirname(__file__), "golden", "dct3d_deblock.npz"))


def vol(name):
    if name == "edge":
        return np.where(np.arange(64)[None, None, :] < 32, 40, 200).astype(np.uint8) * np.ones((64, 64, 1), np.uint8)
    return D[name]


def chunk_of(a, c):
    cz, cy, cx = c // 16, (c // 4) % 4, c % 4
    return a[cz * 16:cz * 16 + 16, cy * 16:cy * 16 + 16, cx * 16:cx * 16 + 16].reshape(-1)


def neighbours(a, c, mask):
    cz, cy, cx = c // 16, (c // 4) % 4, c % 4
    nb = [(cz, cy, cx - 1), (cz, cy, cx + 1), (cz, cy - 1, cx), (cz, cy + 1, cx), (cz - 1, cy, cx), (cz + 1, cy, cx)]
    return [chunk_of(a, (z * 4 + y) * 4 + x) if (mask >> f) & 1 else None for f, (z, y, x) in enumerate(nb)]


def test_volume_api_matches_the_reference():
    for k, name in enumerate(D["case_vol"]):
        v = vol(str(name))
        got = orc.deblock(v, float(D["case_step"][k]), float(D["case_strength"][k]))
        want = D["case_out"][D["case_off"][k]:D["case_off"][k + 1]].reshape(v.shape)
        assert np.array_equal(got, want), (name, D["case_step"][k], D["case_strength"][k])


def test_chunk_api_matches_the_reference():
    a = D["a"]
    for k in range(D["ck_center"].size):
        c, m = int(D["ck_center"][k]), int(D["ck_mask"][k])
        got = orc.deblock_chunk(chunk_of(a, c), neighbours(a, c, m), float(D["ck_step"][k]),
                                float(D["ck_strength"][k]))
        assert np.array_equal(got, D["ck_out"][k]), k


def seam_energy(v, ax):
    """tools/3ddct/test_deblock.c:28-40."""
    d = np.diff(v.astype(np.float64), axis=ax)
    idx = [slice(None)] * 3
    idx[ax] = slice(15, None, 16)  # differences across the planes at 16, 32, 48
    return float(np.mean(d[tuple(idx)] ** 2))


def test_reference_test_checks():
    """tools/3ddct/test_deblock.c:92-137: on the q=24 round trip of the smooth
    ripple, deblocking cuts the seam energy by 20% or more and does not raise
    the error against the original; a genuine step on a block boundary
    survives."""
    z, y, x = np.meshgrid(np.arange(64), np.arange(64), np.arange(64), indexing="ij")
    f = 128.0 + 60.0 * np.sin(x * 0.11 + 0.5) * np.cos(y * 0.09) + 40.0 * np.sin(z * 0.07 + y * 0.03)
    orig = (np.clip(f, 0, 255) + 0.5).astype(np.uint8)
    dec = D["a"]
    deb = orc.deblock(dec, 0.0, 1.0)
    s_pre = sum(seam_energy(dec, ax) for ax in range(3))
    s_post = sum(seam_energy(deb, ax) for ax in range(3))
    assert s_post < 0.8 * s_pre, (s_pre, s_post)
    mse = lambda p: float(np.mean((p.astype(np.float64) - orig) ** 2))  # noqa: E731
    assert mse(deb) <= 1.02 * mse(dec)
    e = orc.deblock(vol("edge"), 0.0, 1.0)
    assert float(np.min(e[:, :, 32].astype(np.float64) - e[:, :, 31])) > 0.9 * 160


def test_pass_order_matters():
    """The reference runs x, then y, then z seams in place: z first (the
    volume transposed) gives other bytes, so the golden vectors pin the order."""
    a = D["a"]
    want = D["case_out"][D["case_off"][1]:D["case_off"][2]].reshape(a.shape)  # a, step 24
    zfirst = orc.deblock(a.transpose(2, 1, 0).copy(), 24.0, 1.0).transpose(2, 1, 0)
    assert not np.array_equal(zfirst, want)
