import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def ctx():
    import torch

    import render3d_amd as r3d

    if not torch.cuda.is_available():
        pytest.fail("GPU test without a GPU")
    c = r3d.Context(0)
    yield c
    c.close()


def pytest_runtest_logstart(nodeid, location):
    # name the case before any GPU step, flushed (a fault can end the process)
    print(f"\n[start] {nodeid}", flush=True)
