"""GPU parity of the plain per-gate form against the CPU oracle, through the C ABI.

Bar: complex128 within 1e-12 max-abs per amplitude (the reference tests'
tolerance, ket/test/runtime_gate_numerical.cpp:80-85); data-movement gates
(X, CNOT, SWAP, local swap) and Z/CZ negations bit-exact; complex64 within
2e-6 of the complex128 oracle on normalised states.
"""
import sys

import numpy as np
import pytest

import brk
import dense as D
from devstate import DeviceState
from gatecases import REFERENCE_CASES, random_cases

pytestmark = pytest.mark.gpu

TOL128 = 1e-12
TOL64 = 2e-6


def note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def run_gpu(case, host, dtype=brk.C128):
    st = DeviceState.from_host(host, dtype)
    try:
        case.gpu(st.ptr, st.size, dtype)
        return st.download()
    finally:
        st.free()


@pytest.mark.parametrize("case", REFERENCE_CASES, ids=lambda c: c.name)
def test_reference_cases(oracle, case):
    note(f"gpu gate {case}")
    s = D.ket_fixture_state(4)
    expected = s.copy()
    case.oracle(oracle, expected)
    got = run_gpu(case, s)
    assert np.max(np.abs(got - expected)) < TOL128


@pytest.mark.parametrize("n", [2, 5, 9, 13, 17])
def test_random_cases_c128(oracle, n):
    rng = np.random.default_rng(1000 + n)
    for case in random_cases(n, 48, seed=2000 + n, max_targets=min(3, n), max_controls=min(3, n - 1)):
        note(f"gpu n={n} {case}")
        s = D.random_state(n, rng)
        expected = s.copy()
        case.oracle(oracle, expected)
        got = run_gpu(case, s)
        err = np.max(np.abs(got - expected))
        if case.kind in ("X", "SWAP", "Z"):
            assert err == 0.0, str(case)
        else:
            assert err < TOL128, (str(case), err)


@pytest.mark.parametrize("n", [6, 14])
def test_random_cases_c64(oracle, n):
    rng = np.random.default_rng(3000 + n)
    for case in random_cases(n, 26, seed=4000 + n):
        if case.kind == "M1":
            continue  # random non-unitary matrices amplify float rounding
        note(f"gpu c64 n={n} {case}")
        s = D.random_state(n, rng)
        expected = s.copy()
        case.oracle(oracle, expected)
        got = run_gpu(case, s.astype(np.complex64), brk.C64).astype(np.complex128)
        assert np.max(np.abs(got - expected)) < TOL64, str(case)


@pytest.mark.parametrize("n,targets", [(20, (0, 19)), (20, (19, 0)), (20, (3, 4)), (20, (0, 1)), (21, (7, 15))])
def test_local_swap_exact(oracle, n, targets):
    note(f"gpu local swap n={n} {targets}")
    rng = np.random.default_rng(n + targets[0])
    s = D.random_state(n, rng)
    expected = s.copy()
    oracle.swap_permutated_local_qubits(expected, *targets)
    st = DeviceState.from_host(s)
    brk.swap_permutated_local_qubits(st.ptr, st.size, brk.C128, *targets)
    got = st.download()
    st.free()
    assert np.array_equal(got, expected)


def test_every_target_position_20q(oracle):
    # each qubit position as target, incl. the low ones sharing cache lines
    n = 20
    rng = np.random.default_rng(77)
    s = D.random_state(n, rng)
    st = DeviceState.from_host(s)
    expected = s.copy()
    for q in range(n):
        note(f"gpu 20q chain q={q}")
        brk.hadamard(st.ptr, st.size, brk.C128, q)
        oracle.hadamard(expected, q)
        brk.phase_shift3(st.ptr, st.size, brk.C128, 0.3 + q, 1.1, -0.7, q)
        oracle.phase_shift3(expected, 0.3 + q, 1.1, -0.7, q)
        brk.exponential_pauli_z(st.ptr, st.size, brk.C128, -0.1 * q, [q])
        oracle.exponential_pauli_z(expected, -0.1 * q, [q])
        brk.controlled_not(st.ptr, st.size, brk.C128, (q + 7) % n, [q])
        oracle.controlled_not(expected, (q + 7) % n, [q])
        brk.pauli_z(st.ptr, st.size, brk.C128, [q, (q + 3) % n])
        oracle.pauli_z(expected, [q, (q + 3) % n])
    got = st.download()
    st.free()
    assert np.max(np.abs(got - expected)) < TOL128


def test_errors_are_reported(oracle):
    st = DeviceState(4)
    with pytest.raises(brk.BrkError) as e:
        brk.hadamard(st.ptr, st.size, brk.C128, 9)
    assert e.value.status == 2
    with pytest.raises(brk.BrkError):
        brk.swap(st.ptr, st.size, brk.C128, 1, 1)
    st.free()


def _band_cases(n):
    """Gates placed on every band the dispatcher distinguishes: lane bits (< 6),
    line bits (< 2 for c128), and high bits, for targets and controls."""
    from gatecases import Case
    out = []
    qs = sorted({0, 1, 2, 3, 5, min(6, n - 1), n - 1})
    for t in qs:
        for ctrl in ([], [c for c in qs if c != t][:1], [c for c in qs if c != t][-1:],
                     [c for c in qs if c != t][:2], [c for c in qs if c != t][-2:]):
            out.append(Case(f"H t{t} c{ctrl}", "H", [t], list(ctrl)))
            out.append(Case(f"U3 t{t} c{ctrl}", "U3", [t], list(ctrl), (0.3, 1.2, -0.8)))
            out.append(Case(f"X t{t} c{ctrl}", "X", [t], list(ctrl)))
            out.append(Case(f"Z t{t} c{ctrl}", "Z", [t], list(ctrl)))
            out.append(Case(f"EZ t{t} c{ctrl}", "EZ", [t], list(ctrl), (0.4,)))
            t2 = [q for q in qs if q != t and q not in ctrl][-1:]
            out.append(Case(f"ZN t{[t] + t2} c{ctrl}", "ZN", [t] + t2, list(ctrl)))
            out.append(Case(f"SZN t{[t] + t2} c{ctrl}", "SZN", [t] + t2, list(ctrl)))
            out.append(Case(f"SZN+ t{t2 + [t]} c{ctrl}", "SZN+", t2 + [t], list(ctrl)))
    for a in qs:
        for b in qs:
            if a < b:
                out.append(Case(f"SWAP {a},{b}", "SWAP", [a, b]))
                out.append(Case(f"SWAP {b},{a}", "SWAP", [b, a]))
                c = [q for q in qs if q not in (a, b)][:1]
                out.append(Case(f"SWAP {a},{b} c{c}", "SWAP", [a, b], c))
    return out


@pytest.mark.parametrize("n", [6, 7, 10])
def test_kernel_bands_c128(oracle, n):
    rng = np.random.default_rng(500 + n)
    for case in _band_cases(n):
        note(f"gpu band n={n} {case}")
        s = D.random_state(n, rng)
        expected = s.copy()
        case.oracle(oracle, expected)
        got = run_gpu(case, s)
        err = np.max(np.abs(got - expected))
        if case.kind in ("X", "SWAP", "Z"):
            assert err == 0.0, str(case)
        else:
            assert err < TOL128, (str(case), err)


@pytest.mark.parametrize("n", [7])
def test_kernel_bands_c64(oracle, n):
    rng = np.random.default_rng(600 + n)
    for case in _band_cases(n):
        note(f"gpu band c64 n={n} {case}")
        s = D.random_state(n, rng)
        expected = s.copy()
        case.oracle(oracle, expected)
        got = run_gpu(case, s.astype(np.complex64), brk.C64).astype(np.complex128)
        assert np.max(np.abs(got - expected)) < TOL64, str(case)


@pytest.mark.parametrize("line_bytes", [1024, 4096])
def test_wide_line_setting_keeps_lane_kernels_in_range(oracle, line_bytes):
    # With lines wider than a wavefront's amplitudes, targets 6.. must not reach the
    # lane kernels (a shuffle mask of 2^6 or more would read the lane's own value).
    n = 12
    rng = np.random.default_rng(line_bytes)
    s = D.random_state(n, rng)
    expected = s.copy()
    st = DeviceState.from_host(s)
    brk.tune("line_bytes", line_bytes)
    try:
        for q in range(10):
            note(f"gpu line_bytes={line_bytes} q={q}")
            brk.hadamard(st.ptr, st.size, brk.C128, q)
            oracle.hadamard(expected, q)
            brk.phase_shift3(st.ptr, st.size, brk.C128, 0.2 * q, 0.5, -1.1, q, [(q + 5) % n])
            oracle.phase_shift3(expected, 0.2 * q, 0.5, -1.1, q, [(q + 5) % n])
            brk.swap(st.ptr, st.size, brk.C128, q, (q + 7) % n)
            oracle.swap(expected, q, (q + 7) % n)
        got = st.download()
    finally:
        brk.tune("line_bytes", 64)
        st.free()
    assert np.max(np.abs(got - expected)) < TOL128
