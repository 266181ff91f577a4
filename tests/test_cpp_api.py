"""The header-only C++20 API (include/gendil_amd/operators.hpp, krylov.hpp):
it builds with a plain C++ compiler against the C-ABI library, and on the GPU
its operators match the oracle and its solvers converge (tests/cpp)."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CPP = os.path.join(HERE, "cpp")
EXE = os.path.join(CPP, "test_operators")


def build():
    res = subprocess.run(["make", "-s", "-C", CPP], capture_output=True, text=True)
    assert res.returncode == 0, res.stdout + res.stderr


def test_cpp_api_builds():
    build()
    assert os.path.exists(EXE)


@pytest.mark.gpu
def test_cpp_api_on_gpu():
    if not os.path.exists(EXE):
        build()
    print("running tests/cpp/test_operators", flush=True)
    res = subprocess.run([EXE], capture_output=True, text=True, timeout=300)
    print(res.stdout)
    assert res.returncode == 0, res.stdout + res.stderr
