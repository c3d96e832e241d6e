"""The C++ mirror header compiles against the C ABI (CPU) and, on the GPU,
the example program recovers a known translation through it."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "mirror_smoke.cpp")
BIN = os.path.join(ROOT, "tests", "cpp", "mirror_smoke")
LIBDIR = os.path.join(ROOT, "sycl-optic-flow_amd", "lib")


def compile_cmd(out):
    return ["g++", "-std=c++17", "-O2", SRC, "-I", os.path.join(ROOT, "include"), "-L", LIBDIR,
            "-lamdflow", f"-Wl,-rpath,{LIBDIR}", "-o", out]


def test_mirror_header_compiles_and_links(tmp_path):
    out = str(tmp_path / "mirror_smoke")
    res = subprocess.run(compile_cmd(out), capture_output=True, text=True)
    assert res.returncode == 0, res.stderr


@pytest.mark.gpu
def test_mirror_program_runs_on_gpu():
    assert os.path.exists(BIN), "tests/cpp/mirror_smoke is built by __graft_entry__.build()"
    res = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    print(res.stdout)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "mirror_smoke OK" in res.stdout
