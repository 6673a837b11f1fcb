"""GPU parity of the FP64 reductions against the CPU oracle (through the C ABI)."""
import sys

import numpy as np
import pytest

import brk
import dense as D
from devstate import DeviceState

pytestmark = pytest.mark.gpu

TOL = 1e-12


def note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_reference_observable_fixtures(oracle):
    note("gpu reductions: reference fixtures")
    p1, p2 = D.observable_state1(4), D.observable_state2(4)
    a, b = DeviceState.from_host(p1), DeviceState.from_host(p2)
    assert abs(brk.inner_product(a.ptr, brk.C128, b.ptr, brk.C128, 16) - oracle.inner_product(p1, p2)) < TOL
    assert abs(brk.norm2(a.ptr, 16, brk.C128) - oracle.norm2(p1)) < TOL
    ip = oracle.inner_product(p1, p2)
    out = (brk.C.c_double * 1)()
    brk.call("brk_fidelity", a.ptr, brk.C128, b.ptr, brk.C128, 16, out, None)
    assert abs(out[0] - abs(ip) ** 2) < 1e-10
    a.free()
    b.free()


@pytest.mark.parametrize("n", [3, 11, 20])
def test_spin_and_pauli(oracle, n):
    rng = np.random.default_rng(n)
    s = D.random_state(n, rng)
    st = DeviceState.from_host(s)
    for q in sorted({0, 1, n // 2, n - 1}):
        note(f"gpu spin n={n} q={q}")
        g = brk.spin_expectation_value(st.ptr, st.size, brk.C128, q)
        o = oracle.spin_expectation_value(s, q)
        assert np.max(np.abs(np.array(g) - np.array(o))) < TOL
    for pauli, qs in (("X", [0]), ("ZZ", [0, n - 1]), ("XYZ", [n - 1, 0, 1]), ("YY", [1, 2])):
        if max(qs) >= n:
            continue
        note(f"gpu pauli n={n} {pauli}{qs}")
        g = brk.expectation_pauli_string(st.ptr, st.size, brk.C128, pauli, qs)
        o = oracle.expectation_pauli_string(s, pauli, qs)
        assert abs(g - o) < TOL
    st.free()


@pytest.mark.parametrize("n,dtype", [(10, brk.C128), (16, brk.C128), (12, brk.C64)])
def test_ising(oracle, n, dtype):
    note(f"gpu ising n={n} dtype={dtype}")
    rng = np.random.default_rng(50 + n)
    s = D.random_state(n, rng)
    J = rng.uniform(-1, 1, n - 1)
    h = rng.uniform(-1, 1, n)
    tol = TOL if dtype == brk.C128 else 1e-6
    host = s if dtype == brk.C128 else s.astype(np.complex64)
    ref = host.astype(np.complex128)
    st = DeviceState.from_host(host, dtype)
    e = brk.expectation_ising(st.ptr, st.size, dtype, J, h)
    assert abs(e - oracle.expectation_ising(ref, J, h)) < tol
    out = DeviceState(n, dtype)
    brk.apply_ising(out.ptr, st.ptr, st.size, dtype, J, h)
    got = out.download().astype(np.complex128)
    assert np.max(np.abs(got - oracle.apply_ising(ref, J, h))) < (tol * 10)
    st.free()
    out.free()


def test_mixed_precision_inner_product(oracle):
    note("gpu mixed inner product")
    rng = np.random.default_rng(9)
    s = D.random_state(15, rng)
    a = DeviceState.from_host(s)
    b = DeviceState.from_host(s.astype(np.complex64), brk.C64)
    ip = brk.inner_product(b.ptr, brk.C64, a.ptr, brk.C128, a.size)
    ref = oracle.inner_product(s.astype(np.complex64).astype(np.complex128), s)
    assert abs(ip - ref) < 1e-12
    a.free()
    b.free()


def test_reductions_are_deterministic():
    note("gpu reduction determinism")
    st = DeviceState(22)
    brk.fill_hash(st.ptr, st.size, brk.C128, 123, 1e-3)
    vals = {brk.norm2(st.ptr, st.size, brk.C128) for _ in range(3)}
    assert len(vals) == 1
    st.free()


def test_fill_hash_matches_numpy_twin():
    note("gpu fill_hash twin")
    st = DeviceState(12)
    brk.fill_hash(st.ptr, st.size, brk.C128, 99, 0.5)
    got = st.download()
    st.free()
    assert np.array_equal(got, D.hash_state_values(99, np.arange(4096), 0.5))
