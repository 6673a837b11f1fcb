"""Pins the CPU oracle's gate functions.

Checks: (1) every reference test case of ket/test/runtime_gate_numerical.cpp on
its fixture state against the closed-form matrix applied by definition
(tests/dense.py); (2) random gates/qubits/controls on random states, n <= 8;
(3) the index-mask construction against brute-force enumeration; (4) invariants
(involutions, unitarity, CZ symmetry) that a sign/index slip would break.
"""
import math
import itertools

import numpy as np
import pytest

import dense as D
from gatecases import REFERENCE_CASES, random_cases

TOL = 1e-12  # ket/test/runtime_gate_numerical.cpp:80-85


def max_err(a, b):
    return float(np.max(np.abs(a - b)))


@pytest.mark.parametrize("case", REFERENCE_CASES, ids=lambda c: c.name)
def test_reference_cases_vs_closed_form(oracle, case):
    s = D.ket_fixture_state(4)
    expected = D.apply(s, *case.dense())
    got = s.copy()
    case.oracle(oracle, got)
    assert max_err(got, expected) < TOL


@pytest.mark.parametrize("n", [1, 2, 3, 5, 8])
def test_random_cases_vs_closed_form(oracle, n):
    rng = np.random.default_rng(n)
    for case in random_cases(n, 60, seed=100 + n, max_targets=min(3, n)):
        s = D.random_state(n, rng)
        expected = D.apply(s, *case.dense())
        got = s.copy()
        case.oracle(oracle, got)
        assert max_err(got, expected) < TOL, str(case)


def test_index_masks_enumerate_each_group_once(oracle):
    # gate_bitmasks.hpp:50-130 + index_with_qubits.hpp:91-112: the 2^(n-k) groups x 2^k local
    # values address every index exactly once, and local value v has bit j = bit of qubits[j].
    n = 7
    for qubits in ([3], [0, 6], [5, 1, 3], [6, 0, 2, 4]):
        k = len(qubits)
        seen = []
        for g in range(2 ** (n - k)):
            for v in range(2 ** k):
                i = oracle.index_with_qubits(g, v, qubits)
                for j, q in enumerate(qubits):
                    assert ((i >> q) & 1) == ((v >> j) & 1)
                seen.append(i)
        assert sorted(seen) == list(range(2 ** n))


def test_hadamard_closed_form_values(oracle):
    # H|0> = (|0>+|1>)/sqrt2 on every qubit -> uniform amplitudes 2^{-n/2}
    # (bra/qcx/hadamards08.qcx applies H to all 8 qubits of |0..0>)
    n = 8
    s = np.zeros(2 ** n, dtype=np.complex128)
    s[0] = 1
    for q in range(n):
        oracle.hadamard(s, q)
    assert max_err(s, np.full(2 ** n, 2 ** (-n / 2))) < TOL


@pytest.mark.parametrize("kind", ["H", "X", "Y", "Z", "SWAP"])
def test_involutions(oracle, kind):
    rng = np.random.default_rng(7)
    n = 6
    s = D.random_state(n, rng)
    t = s.copy()
    for _ in range(2):
        if kind == "H":
            oracle.hadamard(t, 4, [1])
        elif kind == "X":
            oracle.pauli_x(t, [0, 5], [2])
        elif kind == "Y":
            oracle.pauli_y(t, [3], [])
        elif kind == "Z":
            oracle.pauli_z(t, [1, 4])
        else:
            oracle.swap(t, 0, 5, [3])
    assert max_err(s, t) < TOL


def test_u3_inverse_and_norm(oracle):
    rng = np.random.default_rng(3)
    n = 5
    s = D.random_state(n, rng)
    t = s.copy()
    th, ph, la = 1.1, 0.4, -2.3
    oracle.phase_shift3(t, th, ph, la, 2)
    assert abs(oracle.norm2(t) - 1.0) < TOL
    # U3(t,p,l)^dagger = U3(-t,-l,-p)
    oracle.phase_shift3(t, -th, -la, -ph, 2)
    assert max_err(s, t) < TOL


def test_cz_symmetric(oracle):
    rng = np.random.default_rng(4)
    s = D.random_state(6, rng)
    a, b = s.copy(), s.copy()
    oracle.pauli_z(a, [1, 4])
    oracle.pauli_z(b, [4, 1])
    assert max_err(a, b) == 0.0


def test_rz_ry_conventions(oracle):
    # RZ(theta) = EZ(-theta/2), RY(theta) = EY(-theta/2) (exponential_pauli_z.hpp:37-38,
    # exponential_pauli_y.hpp:38)
    rng = np.random.default_rng(5)
    s = D.random_state(4, rng)
    th = 0.77
    a = s.copy()
    oracle.exponential_pauli_z(a, -th / 2, [2])
    assert max_err(a, D.apply(s, D.RZ(th), [2])) < TOL
    b = s.copy()
    oracle.exponential_pauli_y(b, -th / 2, [1])
    assert max_err(b, D.apply(s, D.RY(th), [1])) < TOL


def test_planted_mistake_is_caught(oracle):
    # A transposed U3 (a plausible slip) must fail the closed-form comparison.
    rng = np.random.default_rng(6)
    s = D.random_state(3, rng)
    got = s.copy()
    oracle.phase_shift3(got, 0.9, 0.3, 1.7, 1)
    wrong = D.apply(s, D.U3(0.9, 0.3, 1.7).T, [1])
    assert max_err(got, wrong) > 1e-3
