"""CPU checks of the ket-shaped C++ facade (riken-braket_amd/include/ket_hip.hpp):
it compiles with its own ket::qubit / ket::control and, when the reference's
ket/qubit.hpp and ket/control.hpp are available, with those types instead."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "ket_facade_check.cpp")
INC = ["-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(ROOT, "riken-braket_amd", "include")]
REF_KET = "/root/reference/ket/include"


def compile_only(*extra):
    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("g++ not available")
    cmd = [cxx, "-fsyntax-only", "-std=c++17", "-Wall", "-Wextra", "-Werror", *INC, *extra, SRC]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_facade_compiles_with_own_types():
    compile_only()


@pytest.mark.skipif(not os.path.isfile(os.path.join(REF_KET, "ket", "qubit.hpp")),
                    reason="reference ket headers not present")
def test_facade_compiles_with_reference_types():
    compile_only("-DKET_FACADE_WITH_REFERENCE_TYPES", "-I" + REF_KET)


def test_facade_rejects_bad_call_at_compile_time():
    # a control passed where a phase is expected must not compile
    src = ('#include "ket_hip.hpp"\n'
           'void f(ket::hip::state<std::complex<double>>& s) {\n'
           '  using namespace ket::literals;\n'
           '  ket::gate::ranges::phase_shift(ket::hip::make_policy(), s, 0_cq, 1_q);\n'
           '}\n')
    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("g++ not available")
    r = subprocess.run([cxx, "-fsyntax-only", "-std=c++17", *INC, "-x", "c++", "-"], input=src,
                       capture_output=True, text=True)
    assert r.returncode != 0
