"""CPU check of the streaming Kronecker kernel's global->LDS copy addresses
(gendil_amd/csrc/gd_k3s.hpp): tests/cpp/stream_bounds.hip, compiled for the
host with hipcc, walks every lane of every copy instruction of every tile of
a set of meshes (all periodicities and upwind directions, FP64 and FP32,
advection and mass-only instances) and checks that each source lies in its
buffer and is 16-byte aligned; that the element region's swizzled chunk
layout receives every chunk of every element once, where the kernel's reads
expect it; and that the kernel's tile walk visits every tile once for a range
of tile counts and grids, in both directions.  No GPU is used."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.mark.skipif(not shutil.which(HIPCC) and not os.path.exists(HIPCC), reason="hipcc not installed")
def test_every_stream_copy_source_is_in_bounds(tmp_path):
    exe = tmp_path / "stream_bounds"
    cmd = [HIPCC, "-O2", "-std=c++20", "--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang",
           "-packed-fp32-ops", f"-I{ROOT}/include", f"-I{ROOT}/gendil_amd/csrc",
           f"{ROOT}/tests/cpp/stream_bounds.hip", "-o", str(exe)]
    subprocess.run(cmd, check=True, capture_output=True, timeout=300)
    res = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert res.returncode == 0, res.stdout[-2000:]
    assert "TOTAL BAD 0" in res.stdout
