"""Host-side pieces of bench.py: the FP64 decode used to report the MFMA
variants' error, checked against the oracle's FP32 decode before rounding."""
import numpy as np
import pytest

import bench
from oracle import oracle as orc

torch = pytest.importorskip("torch")


def test_fp64_decode_matches_oracle_pre_round():
    rng = np.random.default_rng(9)
    n = 6
    lv = (rng.integers(-300, 300, (n, 4096)) * (rng.random((n, 4096)) < 0.2)).astype(np.int16)
    hdr = orc.hdr_array(rng.uniform(-10, 20, n), rng.uniform(1, 255, n), rng.integers(-500, 3000, n))
    step = orc.step_linear()
    got = bench.fp64_decode(torch.from_numpy(lv), torch.from_numpy(hdr), torch.from_numpy(step)).numpy()
    for i in range(n):
        ok, _, pre = orc.decode_chunk(lv[i], hdr[i, 0:1].view(np.float32)[0], hdr[i, 1:2].view(np.float32)[0],
                                      hdr[i, 2], step)
        assert ok
        err = np.abs(got[i] - pre.astype(np.float64))
        assert err.max() < 2e-3 * max(1.0, np.abs(pre).max() / 255), err.max()


def test_multi_gpu_request_exits_with_message(monkeypatch, capsys):
    monkeypatch.setattr("sys.argv", ["bench.py", "--gpus", "2"])
    assert bench.main() == 2
    assert "one GPU" in capsys.readouterr().err
