"""The C-ABI library loads, exports every entry point include/r3d_hip.h
declares, agrees with the ctypes mirror, and compiles as C23 (no GPU calls
that compute)."""
import ctypes as C
import os
import re
import subprocess

import pytest

from render3d_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "r3d_hip.h")


def declared():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"\b(r3d_hip_\w+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(_lib.LIB_PATH)
    names = declared()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_ctypes_mirror_covers_the_header():
    assert sorted(_lib.SIGNATURES) == declared()
    _lib.load()


def test_struct_sizes_match_the_c_compiler(tmp_path):
    src = tmp_path / "sz.c"
    src.write_text('#include "r3d_hip.h"\n#include <stdio.h>\n#include <stddef.h>\nint main(void){printf("%zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu\\n",'
                   'sizeof(r3d_frame_params), sizeof(r3d_chunk_batch), sizeof(r3d_chunk_header), sizeof(r3d_hip_image),'
                   'sizeof(r3d_raycast_opts), sizeof(r3d_diff_params), offsetof(r3d_chunk_batch, out), '
                   'sizeof(r3d_raycast_tuning), sizeof(r3d_blob_batch), sizeof(r3d_blob_parse_out), '
                   'sizeof(r3d_shard_src), sizeof(r3d_shard_region), sizeof(r3d_bricks_desc));return 0;}\n')
    exe = tmp_path / "sz"
    subprocess.run(["gcc", "-std=c2x", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    want = [C.sizeof(_lib.FrameParams), C.sizeof(_lib.ChunkBatch), C.sizeof(_lib.ChunkHeader), C.sizeof(_lib.Image),
            C.sizeof(_lib.RaycastOpts), C.sizeof(_lib.DiffParams), _lib.ChunkBatch.out.offset,
            C.sizeof(_lib.RaycastTuning), C.sizeof(_lib.BlobBatch), C.sizeof(_lib.BlobParseOut),
            C.sizeof(_lib.ShardSrc), C.sizeof(_lib.ShardRegion), C.sizeof(_lib.BricksDesc)]
    assert got == want


@pytest.mark.skipif(os.environ.get("HIP_VISIBLE_DEVICES") is None and os.path.exists("/dev/kfd"),
                    reason="the no-device failure path is checked where no GPU is present")
def test_c23_program_links_and_fails_cleanly_without_a_gpu(tmp_path):
    exe = tmp_path / "abi"
    subprocess.run(["gcc", "-std=c2x", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "c", "abi_check.c"), "-o", str(exe),
                    "-L", os.path.dirname(_lib.LIB_PATH), "-lr3d_hip",
                    "-Wl,-rpath," + os.path.dirname(_lib.LIB_PATH)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, (r.returncode, r.stdout, r.stderr)
