"""Host-only checks of the serving loop's building blocks (no GPU).

* The budget allocator (mublis_amd/runtime/budget_arena.h) against a
  brute-force model of every 256-byte unit, built with AddressSanitizer and
  UndefinedBehaviorSanitizer: no overlap, best fit never misses a free run
  large enough, coalescing leaves no two adjacent free blocks, double release
  is refused.
(The serving structs' layout is checked with the other public structs in
test_abi_layout.py.)
"""
import os
import shutil
import subprocess

import pytest

import l3cases as L

HERE = os.path.dirname(os.path.abspath(__file__))


def _gxx():
    gxx = shutil.which("g++")
    if not gxx:
        pytest.skip("g++ not available")
    return gxx


def test_budget_arena_under_sanitizers(tmp_path):
    exe = tmp_path / "arena_check"
    subprocess.run([_gxx(), "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined",
                    "-fno-sanitize-recover=all",
                    "-I", os.path.join(L.ROOT, "mublis_amd", "runtime"),
                    os.path.join(HERE, "native", "arena_check.cpp"), "-o", str(exe)],
                   check=True)
    for seed in (1, 2, 3):
        out = subprocess.run([str(exe), str(seed)], capture_output=True, text=True,
                             timeout=300)
        assert out.returncode == 0, out.stderr
        assert f"ok seed {seed}" in out.stdout
