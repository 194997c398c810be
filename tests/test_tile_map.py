"""The tiled stencil's thread -> site mapping, checked on the CPU.

Compiles tests/tile_map_check.cpp against minim_amd/csrc/pf_tile.hpp (the
same header the HIP kernel uses) with the host compiler and runs it: every
site owned exactly once, every load inside the grid, every neighbour read
from LDS or registers equal to the periodic neighbour.
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def test_tile_mapping(tmp_path):
    cxx = shutil.which("g++") or shutil.which("c++")
    if cxx is None:
        pytest.skip("no host C++ compiler")
    exe = str(tmp_path / "tile_map_check")
    subprocess.run([cxx, "-O2", "-std=c++17", "-I", os.path.join(ROOT, "minim_amd", "csrc"),
                    os.path.join(HERE, "tile_map_check.cpp"), "-o", exe], check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    print(r.stdout[-2000:])
    assert r.returncode == 0, r.stdout[-2000:]
