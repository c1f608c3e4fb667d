"""Host-only check of the GEMM kernels' workgroup -> tile maps (no GPU).

tests/native/walk_check.cpp includes mublis_amd/targets/gfx950/gemm_tile.h,
the header the kernels are built from, compiled for the host only:
map_tile's full and triangle grids (including the SYRK split into 128-tiles
strictly inside the triangle plus diagonal blocks as 32-tiles, which must
cover every 32-tile of the triangle exactly once), and map_tile_lpt's
long-first walk (a bijection onto the tiles; every XCD gets the same total
K up to a bound that round-robin dealing would exceed; each XCD meets its
tiles in non-increasing K).
"""
import os
import shutil
import subprocess

import pytest

import l3cases as L

HERE = os.path.dirname(os.path.abspath(__file__))


def test_tile_walks(tmp_path):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    exe = tmp_path / "walk_check"
    subprocess.run([hipcc, "-x", "hip", "--cuda-host-only", "-O1", "-std=c++17",
                    "-I", os.path.join(L.ROOT, "mublis_amd", "targets", "gfx950"),
                    os.path.join(HERE, "native", "walk_check.cpp"), "-o", str(exe)],
                   check=True, cwd=tmp_path)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.strip().endswith("ok")


def test_bf16_split(tmp_path):
    """tests/native/split_check.cpp: the BF16x3 split (bf16_split.h) is exact
    on a dense sample of all FP32 bit patterns and on the whole top binade,
    keeps Inf and NaN in the high part, and a planted split without the
    overflow guard fails the same check."""
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    exe = tmp_path / "split_check"
    subprocess.run([hipcc, "-x", "hip", "--cuda-host-only", "-O2", "-std=c++17",
                    "-I", os.path.join(L.ROOT, "mublis_amd", "targets", "gfx950"),
                    os.path.join(HERE, "native", "split_check.cpp"), "-o", str(exe)],
                   check=True, cwd=tmp_path)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.strip().endswith("ok")
