"""The C++ BatchAnalyzer-shaped wrapper (include/hwpa_batch_analyzer.hpp): it compiles
as plain host C++17 against the C-ABI, and on a GPU its BatchResults equal the C-ABI
analyzer's outputs (tests/cpp/wrapper_main.cpp)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _build(tmp_path):
    exe = str(tmp_path / "wrapper_main")
    gxx = shutil.which("g++") or shutil.which("c++")
    if gxx is None:
        pytest.skip("no host C++ compiler")
    cmd = [gxx, "-std=c++17", "-O1", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "cpp", "wrapper_main.cpp"), "-L", os.path.join(ROOT, "hw_amd"), "-lhwpa",
           "-Wl,-rpath," + os.path.join(ROOT, "hw_amd"), "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


def test_wrapper_compiles(tmp_path):
    assert os.path.exists(_build(tmp_path))


@pytest.mark.gpu
def test_wrapper_results_equal_c_abi(tmp_path, gpu):
    exe = _build(tmp_path)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stderr
    assert "wrapper ok" in r.stdout
