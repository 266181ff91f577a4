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


# ---- the GenDiL-side binding (include/gendil_amd/gendil_policy.hpp) ---------

GENDIL_SRC = "/root/reference"
BINDING = os.path.join(CPP, "gendil_binding")


def _have_gendil():
    return os.path.exists(os.path.join(GENDIL_SRC, "include", "gendil", "gendil.hpp"))


@pytest.mark.skipif(not _have_gendil(), reason="needs the GenDiL sources (build container only)")
def test_gendil_binding_compiles_against_gendil():
    # its static_asserts check that the policy-tagged factories return the
    # MI355X operator for every value category of a descriptor, that lambdas
    # still build GenDiL's own operators under GenDiL's policies, and that the
    # solver calls on the MI355X operator resolve to the device-resident loops
    res = subprocess.run(["make", "-s", "-C", CPP, "gendil_binding"], capture_output=True, text=True)
    assert res.returncode == 0, res.stdout + res.stderr
    assert os.path.exists(BINDING)


@pytest.mark.skipif(not _have_gendil(), reason="needs the GenDiL sources (build container only)")
def test_gendil_binding_rejects_lambdas_under_the_mi355x_policy(tmp_path):
    src = tmp_path / "lambda.cpp"
    src.write_text(
        "#include <gendil/gendil.hpp>\n"
        '#include "gendil_amd/gendil_policy.hpp"\n'
        "using namespace gendil;\n"
        "int main() {\n"
        "  CartesianMesh<3> mesh({2, 2, 2}, {0.5, 0.5, 0.5}, {0.0, 0.0, 0.0}, true);\n"
        "  auto fes = MakeFiniteElementSpace(mesh, MakeLobattoFiniteElement(FiniteElementOrders<1, 1, 1>{}));\n"
        "  auto ir = MakeIntegrationRule(IntegrationRuleNumPoints<2, 2, 2>{});\n"
        "  auto adv = [](const std::array<Real, 3>&, Real (&v)[3]) { v[0] = v[1] = v[2] = 1.0; };\n"
        "  auto A = MakeAdvectionOperator<Mi355xKernelConfiguration>(fes, ir, adv);\n"
        "  (void)A;\n"
        "}\n")
    root = os.path.dirname(HERE)
    res = subprocess.run(["g++", "-std=c++20", "-fsyntax-only", "-fopenmp", f"-I{GENDIL_SRC}/include",
                          f"-I{root}/include", str(src)], capture_output=True, text=True)
    assert res.returncode != 0
    assert "device lambdas cannot cross the C-ABI" in res.stderr


@pytest.mark.gpu
def test_gendil_binding_matches_gendil_on_gpu():
    # built in the build container with GenDiL's headers; it must be here
    assert os.path.exists(BINDING), "tests/cpp/gendil_binding missing: build it with `make -C tests/cpp gendil_binding`"
    print("running tests/cpp/gendil_binding", flush=True)
    res = subprocess.run([BINDING], capture_output=True, text=True, timeout=300)
    print(res.stdout)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "all binding checks passed" in res.stdout
