"""GPU check of the ket-shaped C++ facade: tests/cpp/ket_facade_check drives
ket::gate::* (with adj_* and runtime:: qubit-range forms), ket::gate::ranges::*,
ket::mpi::qubit_permutation, the reductions, measurement / sampling with
std::mt19937_64 and the Shor box through riken-braket_amd/include/ket_hip.hpp;
the same sequence is replayed here on the CPU oracle."""
import math
import os
import subprocess
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "ket_facade_check")


def oracle_sequence(o):
    s = np.zeros(256, dtype=np.complex128)
    s[0] = 1
    o.hadamard(s, 0)
    o.hadamard(s, 5)
    o.hadamard(s, 3, [0])
    o.pauli_x(s, [7])
    o.pauli_x(s, [2], [5])
    o.pauli_x(s, [1, 4], [3])
    o.pauli_x(s, [6])
    o.pauli_y(s, [6])
    o.pauli_y(s, [1, 2], [0])
    o.pauli_z(s, [4, 6])
    o.pauli_z(s, [2])
    o.phase_shift(s, 0.3, [1])
    o.phase_shift(s, 0.7, [3, 0])
    o.phase_shift_coeff(s, complex(math.cos(0.5), math.sin(0.5)), [5])
    o.phase_shift2(s, 0.2, 0.9, 6)
    o.phase_shift3(s, 0.1, 0.4, 1.3, 7, [2])
    o.exponential_pauli_x(s, 0.35, [0])
    o.exponential_pauli_y(s, 0.25, [3, 6])
    o.exponential_pauli_z(s, 0.45, [1, 5], [7])
    o.swap(s, 2, 6)
    o.swap(s, 0, 7, [4])
    o.exponential_swap(s, 0.55, 3, 5)
    o.exponential_swap(s, -0.8, 1, 6, [2])
    # adjoint forms and the remaining gates (matrices of the ket headers' comments)
    r = 1 / math.sqrt(2)
    a, b = (1 + 1j) / 2, (1 - 1j) / 2
    e = complex(math.cos(0.25), math.sin(0.25))
    o.phase_shift3(s, -0.1, -1.3, -0.4, 7, [2])           # U3(t, p, l)^+ = U3(-t, -l, -p)
    o.gate_1q(s, 4, [a, b, b, a])                         # sqrt(X)
    o.gate_1q(s, 2, [b, b, -b, b], [5])                   # sqrt(Y)^+
    o.phase_shift_coeff(s, 1j, [1, 3])                    # sqrt(Z) on control qubits
    o.pauli_zn(s, [1, 3])                                 # ZZ on ket::qubit targets
    o.pauli_zn(s, [0, 5, 6], [2])
    o.sqrt_pauli_zn(s, [1, 3])
    o.sqrt_pauli_zn(s, [0, 4, 7], [2], adjoint=True)
    o.phase_shift_coeff(s, 1j, [6, 5])                    # CsZ with a target: i on |11>
    o.gate_1q(s, 6, [r, 1j * r, 1j * r, r])               # +X
    o.gate_1q(s, 0, [r, -r, r, r], [7])                   # (+Y)^+ = -Y
    o.gate_1q(s, 3, [(1 + e) / 2, (1 - e) / 2, (1 - e) / 2, (1 + e) / 2], [0, 1])  # controlled V
    o.pauli_x(s, [5], [2, 6])                             # Toffoli
    o.phase_shift(s, 0.9, [2, 4])                         # controlled phase shift
    o.exponential_pauli_y(s, 0.3, [0, 4])
    o.exponential_swap(s, -0.4, 2, 7)
    o.pauli_x(s, [1, 6], [3])                             # runtime forms
    o.hadamard(s, 2, [3])
    o.exponential_pauli_z(s, 0.7, [1, 6])
    o.pauli_x(s, [1, 6], [0, 5])                          # iterator-pair runtime forms
    o.hadamard(s, 6, [3])
    o.exponential_pauli_y(s, -0.45, [4, 7])
    o.phase_shift3(s, 0.3, -0.2, 0.9, 5)
    o.sqrt_pauli_zn(s, [4, 7], [3])
    perm, inv = list(range(8)), list(range(8))
    o.swap_permutated_local_qubits(s, 1, 6)
    o.permutate(perm, inv, inv[1], inv[6])
    o.hadamard(s, perm[1])
    o.exponential_pauli_x(s, 0.6, [perm[6]], [perm[3]])
    return s, perm


def oracle_measured(o):
    """The measured / sampled second state: bra's generator std::mt19937_64(7)."""
    m = np.zeros(256, dtype=np.complex128)
    m[0] = 1
    for q in range(8):
        o.hadamard(m, q)
    o.phase_shift3(m, 0.7, 0.2, -0.5, 3)
    o.pauli_x(m, [1], [3])
    g = o.mt64(7)
    outcome = o.projective_measurement(m, 3, g.canonical())
    events = o.generate_events(m, [g.canonical() for _ in range(12)]).tolist()
    for q, value in ((5, 0), (6, 1)):
        o.project(m, q, value, o.zero_one_probabilities(m, q)[value] ** -0.5)
    return m, outcome, events


def parse(stdout):
    out = {}
    for line in stdout.splitlines():
        tag, key, *vals = line.split()
        out[(tag, key)] = vals
    return out


def test_facade_program_matches_oracle(oracle, tmp_path):
    if not os.path.isfile(EXE):
        pytest.fail(f"{EXE} is not built (run __graft_entry__.build())")
    print("gpu ket facade: running tests/cpp/ket_facade_check", flush=True)
    sys.stdout.flush()
    r = subprocess.run([EXE, str(tmp_path / "state")], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = parse(r.stdout)
    expected, perm = oracle_sequence(oracle)
    t = np.zeros(256, dtype=np.complex128)
    t[0] = 1
    oracle.hadamard(t, 2)
    ip = oracle.inner_product(expected, t)
    spin = oracle.spin_expectation_value(expected, 4)
    for tag, tol in (("c128", 1e-12), ("c64", 2e-6)):
        state = np.fromfile(tmp_path / f"state_{tag}.bin", dtype=np.complex128)
        assert state.size == 256
        assert np.max(np.abs(state - expected)) < tol, tag
        assert [int(v) for v in got[(tag, "perm")]] == perm
        assert [int(v) for v in got[(tag, "permutate_bits")]] == [
            oracle.permutate_bits(perm, 0x82), oracle.inverse_permutate_bits(perm, [int(x) for x in np.argsort(perm)], 0x82)]
        assert np.allclose([float(v) for v in got[(tag, "spin4")]], spin, atol=tol, rtol=0)
        gi = complex(*[float(v) for v in got[(tag, "inner")]])
        assert abs(gi - ip) < tol
        assert abs(float(got[(tag, "fidelity")][0]) - abs(ip) ** 2) < tol
        assert got[(tag, "error")] == ["2"]  # BRK_ERR_QUBIT_OUT_OF_RANGE for 9_q on 8 qubits
        # measurement, sampling, CLEAR / SET (same draws), all spin expectation values, Shor box
        m, outcome, events = oracle_measured(oracle)
        assert int(got[(tag, "outcome")][0]) == outcome
        assert [int(v) for v in got[(tag, "events")]] == events
        measured = np.fromfile(tmp_path / f"state_measured_{tag}.bin", dtype=np.complex128)
        assert np.max(np.abs(measured - m)) < tol, tag
        assert np.allclose([float(v) for v in got[(tag, "spins7")]], oracle.spin_expectation_value(m, 7),
                           atol=tol, rtol=0)
        box = np.zeros(256, dtype=np.complex128)
        oracle.shor_box(box, 4, 15, 7)
        assert np.max(np.abs(np.fromfile(tmp_path / f"state_shor_{tag}.bin", dtype=np.complex128) - box)) < tol
