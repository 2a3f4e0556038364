"""A C23 host program drives the GPU path through include/r3d_hip.h alone
(tests/c/gpu_golden.c): it decodes the reference's 40 golden 3ddct blobs
(dct3d_decode_u8's bytes, tests/golden/dct3d_golden.npz) and raycasts a
small two-view scene whose expected image comes from the oracle, and exits
0 only when its bytes match; the files it writes are compared here again."""
import os
import subprocess

import numpy as np
import pytest

import render3d_amd as r3d
from oracle import oracle as orc
from render3d_amd import _lib, scene

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
GOLD = np.load(os.path.join(HERE, "golden", "dct3d_golden.npz"))


def _write(d, name, arr):
    np.ascontiguousarray(arr).tofile(os.path.join(d, name))


def test_c23_program_decodes_and_raycasts(tmp_path):
    d = str(tmp_path)
    offs = GOLD["blob_offsets"]
    _write(d, "blobs.bin", GOLD["blob_bytes"])
    _write(d, "offsets.bin", offs[:-1].astype(np.uint64))
    _write(d, "lengths.bin", np.diff(offs).astype(np.uint32))
    _write(d, "expect_u8.bin", GOLD["blob_u8"])

    nx, ny, nz, W, H = 56, 48, 40, 96, 64
    vol = orc.synth_sheets(nx, ny, nz, 3, 5)
    lut = scene.tf_ramp_40_200()
    views = [scene.frame_params(scene.orbit_camera(35.0, 20.0, 2.0), W // 2, H, view_org=(0, 0), step_voxels=0.75),
             scene.frame_params(scene.orbit_camera(160.0, -15.0, 1.8), W // 2, H, view_org=(W // 2, 0),
                                mode=r3d.MODE_MIP)]
    head = np.array([nx, ny, nz, W, H, len(views)], np.uint32).tobytes()
    vbytes = b"".join(bytes(v) for v in views)
    with open(os.path.join(d, "scene.bin"), "wb") as f:
        f.write(head + vbytes + np.ascontiguousarray(lut, np.uint8).tobytes() + vol.tobytes())
    ov = orc.OracleVolume(vol, nx, ny, nz)
    ref, cnt, _ = orc.raycast(ov, scene.tf_unorm(lut), views, W, H, alpha_cutoff=0.99)
    ref8 = (np.clip(ref, 0, 1) * np.float32(255) + np.float32(0.5)).astype(np.uint8)
    ref8[ref[..., 3] == 0] = 0
    _write(d, "expect_rgba8.bin", ref8)
    _write(d, "expect_rgba.bin", ref.astype(np.float32))

    exe = os.path.join(d, "gpu_golden")
    libdir = os.path.dirname(_lib.LIB_PATH)
    subprocess.run(["gcc", "-std=c2x", "-Wall", "-Werror", "-D__HIP_PLATFORM_AMD__", "-I", os.path.join(ROOT, "include"),
                    "-I", "/opt/rocm/include", os.path.join(HERE, "c", "gpu_golden.c"), "-o", exe, "-L", libdir,
                    "-lr3d_hip", "-L", "/opt/rocm/lib", "-lamdhip64", "-lm", "-Wl,-rpath," + libdir,
                    "-Wl,-rpath,/opt/rocm/lib"], check=True)
    print("running the C23 program", flush=True)
    r = subprocess.run([exe, d], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, (r.returncode, r.stdout, r.stderr)
    got = np.fromfile(os.path.join(d, "got_u8.bin"), np.uint8).reshape(-1, 4096)
    assert np.array_equal(got, GOLD["blob_u8"])
    got8 = np.fromfile(os.path.join(d, "got_rgba8.bin"), np.uint8).reshape(H, W, 4)
    assert np.array_equal(got8, ref8)
    gotf = np.fromfile(os.path.join(d, "got_rgba.bin"), np.float32).reshape(H, W, 4)
    assert float(np.max(np.abs(gotf - ref))) <= 1e-5
    assert cnt["samples"] > 1000
