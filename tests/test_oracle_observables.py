"""Pins the oracle's reductions and the qubit permutation / local swap.

Reductions: the brute-force sums of ket/test/runtime_observable_numerical.cpp:101-119
on its fixture states, dense Pauli matrices, and closed forms on product states.
Permutation: the invariant permuted[p] == logical[inverse_permutate_bits(perm, p)]
of ket/test/mpi/runtime_gate_numerical.cpp:111-127, and brute-force bit moves.
"""
import math
import random

import numpy as np
import pytest

import dense as D

TOL = 1e-12


def test_reference_observable_fixtures(oracle):
    # runtime_observable_numerical.cpp:101-119,155-235: <psi1|psi1>, <psi2|psi1>, |<psi2|psi1>|^2
    p1, p2 = D.observable_state1(4), D.observable_state2(4)
    ev = np.sum(np.conj(p1) * p1)
    ip = np.sum(np.conj(p2) * p1)
    assert abs(oracle.inner_product(p1, p1) - ev) < TOL
    assert abs(oracle.inner_product(p1, p2) - ip) < TOL
    assert abs(oracle.norm2(p1) - ev.real) < TOL
    # the reference's identity observable over qubits {0,2,3} gives <psi|psi>
    assert abs(oracle.expectation_pauli_string(p1, "III", [0, 2, 3]) - ev) < TOL


def _pauli_dense(pauli, qubits, n):
    mats = {"I": D.I2, "X": D.X, "Y": D.Y, "Z": D.Z}
    U = D.kron_all([mats[c] for c in pauli])
    return D.full_matrix(n, U, qubits)


@pytest.mark.parametrize("pauli,qubits", [("X", [2]), ("Y", [0]), ("Z", [4]), ("ZZ", [1, 2]),
                                          ("XYZ", [4, 0, 2]), ("YY", [3, 1]), ("IXI", [0, 1, 2])])
def test_pauli_string_vs_dense(oracle, pauli, qubits):
    rng = np.random.default_rng(11)
    n = 5
    s = D.random_state(n, rng)
    M = _pauli_dense(pauli, qubits, n)
    expected = np.vdot(s, M @ s)
    assert abs(oracle.expectation_pauli_string(s, pauli, qubits) - expected) < TOL


def test_spin_expectation_closed_form(oracle):
    # qubit in cos(a)|0> + e^{ib} sin(a)|1>: x = Re(conj a0 a1) = cos a sin a cos b,
    # y = cos a sin a sin b, z = (cos^2 a - sin^2 a)/2 (spin_expectation_value.hpp:47-76)
    a, b = 0.3, 1.2
    q = np.array([math.cos(a), np.exp(1j * b) * math.sin(a)])
    rest = D.random_state(3, np.random.default_rng(2))
    s = np.kron(rest, q).astype(np.complex128)  # qubit 0 = q
    x, y, z = oracle.spin_expectation_value(s, 0)
    assert abs(x - math.cos(a) * math.sin(a) * math.cos(b)) < TOL
    assert abs(y - math.cos(a) * math.sin(a) * math.sin(b)) < TOL
    assert abs(z - 0.5 * (math.cos(a) ** 2 - math.sin(a) ** 2)) < TOL


def test_ising_energy_and_apply(oracle):
    rng = np.random.default_rng(12)
    n = 6
    s = D.random_state(n, rng)
    J = rng.uniform(-1, 1, n - 1)
    h = rng.uniform(-1, 1, n)
    Hm = np.zeros((2 ** n, 2 ** n), dtype=np.complex128)
    for i in range(n - 1):
        Hm += J[i] * _pauli_dense("ZZ", [i, i + 1], n)
    for i in range(n):
        Hm += h[i] * _pauli_dense("X", [i], n)
    e = np.vdot(s, Hm @ s).real
    assert abs(oracle.expectation_ising(s, J, h) - e) < TOL
    assert np.max(np.abs(oracle.apply_ising(s, J, h) - Hm @ s)) < TOL
    # product state |+>^n: <X_i> = 1, <Z_i Z_j> = 0
    plus = np.full(2 ** n, 2 ** (-n / 2), dtype=np.complex128)
    assert abs(oracle.expectation_ising(plus, J, h) - h.sum()) < TOL


def _brute_permutate_bits(perm, x):
    # permutate_bits moves the bit at position b to position perm[b]
    y = 0
    for b, p in enumerate(perm):
        if (x >> b) & 1:
            y |= 1 << p
    return y


def test_permutate_bits_bruteforce(oracle):
    rng = random.Random(5)
    for n in (1, 3, 6, 10):
        for _ in range(10):
            perm = list(range(n))
            rng.shuffle(perm)
            inv = [0] * n
            for q, p in enumerate(perm):
                inv[p] = q
            for _ in range(20):
                x = rng.randrange(2 ** n)
                y = oracle.permutate_bits(perm, x)
                assert y == _brute_permutate_bits(perm, x)
                assert oracle.inverse_permutate_bits(perm, inv, y) == x


def test_permutate_updates_inverse(oracle):
    n = 8
    perm, inv = list(range(n)), list(range(n))
    rng = random.Random(9)
    for _ in range(30):
        a, b = rng.sample(range(n), 2)
        oracle.permutate(perm, inv, a, b)
        assert all(perm[inv[p]] == p for p in range(n))


@pytest.mark.parametrize("p1,p2", [(0, 5), (5, 0), (2, 3), (1, 6), (0, 1)])
def test_local_swap_equals_swap_gate(oracle, p1, p2):
    rng = np.random.default_rng(p1 * 7 + p2)
    n = 7
    s = D.random_state(n, rng)
    a = s.copy()
    oracle.swap_permutated_local_qubits(a, p1, p2)
    b = D.apply(s, D.SWAP, [p1, p2])
    assert np.max(np.abs(a - b)) == 0.0


def test_permutation_invariant_after_local_swaps(oracle):
    # ket/test/mpi/runtime_gate_numerical.cpp:111-127: after data swaps + permutate,
    # physical[p] == logical[inverse_permutate_bits(perm, p)]
    rng = random.Random(21)
    n = 7
    logical = D.random_state(n, np.random.default_rng(3))
    physical = logical.copy()
    perm, inv = list(range(n)), list(range(n))
    for _ in range(6):
        ph, pl = rng.sample(range(n), 2)
        oracle.swap_permutated_local_qubits(physical, ph, pl)
        oracle.permutate(perm, inv, inv[ph], inv[pl])
    for p in range(2 ** n):
        assert physical[p] == logical[oracle.inverse_permutate_bits(perm, inv, p)]
