"""The C++ API as a drop-in: a program written against the reference's class
names and setters (tests/cpp_api_example.cpp) compiles against include/minim/,
links with libminim_amd.so, and runs its host-only part (potential set-up)
without a GPU; the GPU test runs its minimisation too."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "minim_amd")


def build(tmp_path):
    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("g++ not available")
    if not os.path.exists(os.path.join(LIB, "libminim_amd.so")):
        pytest.fail("libminim_amd.so is not built")
    exe = str(tmp_path / "cpp_api_example")
    cmd = [cxx, "-std=c++17", "-O1", os.path.join(ROOT, "tests", "cpp_api_example.cpp"), "-I",
           os.path.join(ROOT, "include"), "-L", LIB, "-lminim_amd", "-Wl,-rpath," + LIB, "-o", exe]
    subprocess.run(cmd, check=True, capture_output=True, text=True, timeout=300)
    return exe


def test_cpp_api_compiles_links_and_sets_up(tmp_path):
    exe = build(tmp_path)
    r = subprocess.run([exe, "host"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ok" in r.stdout


@pytest.mark.gpu
def test_cpp_api_minimises_on_the_gpu(gpu, tmp_path):
    exe = build(tmp_path)
    r = subprocess.run([exe, "gpu"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
