"""The C++ IComputable mirror (host/computable.hpp) and the reference's MSTest
suites restated in C++ (tests/cpp/icomputable_tests.cpp)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "gpgpu-binary-automaton-synchronizing-algorithms_amd")
ORACLE = os.path.join(ROOT, "oracle")


def build(out):
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-I", os.path.join(ROOT, "include"), "-I", os.path.join(PKG, "host"),
           os.path.join(ROOT, "tests", "cpp", "icomputable_tests.cpp"), "-o", out,
           "-L", PKG, "-lssw", "-L", ORACLE, "-lssw_oracle", f"-Wl,-rpath,{PKG}", f"-Wl,-rpath,{ORACLE}"]
    subprocess.run(cmd, check=True, capture_output=True, text=True)


def test_cpp_host_compiles(tmp_path):
    build(str(tmp_path / "icomputable_tests"))


@pytest.mark.gpu
def test_cpp_icomputable_suites(tmp_path):
    exe = str(tmp_path / "icomputable_tests")
    build(exe)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-3000:]
    assert "passed" in r.stdout
