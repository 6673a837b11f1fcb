"""Whole-circuit parity at 30 qubits (16 GiB complex<double>): the 208-op random
circuit of seed 2 (200 gates -- H, X, RZ, U3 on random targets, CNOT and CZ on
random pairs -- plus 8 local swaps), the circuit the 32-qubit timing runs.  A
state this size takes the non-temporal kernel instantiations (states of 256 MiB
and more), so these are the kernels the timing measures.

Start state: the device hash fill (brk_fill_hash), copied to the host for the
CPU oracle.  Checks:
* per-gate run (brk_circuit_run) vs the oracle: max-abs <= 1e-12 over every
  amplitude, the same final permutation
  (ket/test/runtime_gate_numerical.cpp:80-85 tolerance);
* gate blocks (brk_circuit_run_blocked): bit-identical to the per-gate run;
* precision form (fused complex64 on MFMA) at widths 4, 5, 6: infidelity
  1 - |<ref|psi>|^2 <= 1e-6 and max-abs <= 2e-6 on normalised states vs the
  oracle, the same final permutation.
"""
import sys

import numpy as np
import pytest

import brk
import circuits as W
import hip
from devstate import DeviceState
from oracle_circuits import run_ops

pytestmark = [pytest.mark.gpu, pytest.mark.slow, pytest.mark.timeout(900)]

N = 30
SEED = 2
FILL_SEED = 4321
SCALE = 2.0 ** -15
CHUNK = 1 << 26  # amplitudes compared per step (1 GiB of complex128)


def note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _ops():
    ops = W.random_circuit(N, 200, 8, seed=SEED)
    assert len(ops) == 208
    return ops


def _download_range(st, start, count):
    b = 16 if st.dtype == brk.C128 else 8
    out = np.empty(count, dtype=np.complex128 if st.dtype == brk.C128 else np.complex64)
    brk.call("brk_memcpy_dtoh", out.ctypes.data, st.ptr + start * b, count * b, None)
    return out


@pytest.fixture(scope="module")
def reference():
    free, _ = hip.mem_info()
    if free < (26 << 30):
        pytest.skip("needs 26 GiB of free device memory")
    note("30q: hash fill, copy to host")
    st = DeviceState(N)
    brk.fill_hash(st.ptr, st.size, brk.C128, FILL_SEED, SCALE)
    expected = st.download()
    st.free()
    note("30q: oracle run of the 208 ops")
    from oracle import Oracle
    perm, done = run_ops(Oracle(), expected, _ops())
    assert done == 208
    note("30q: oracle done")
    return expected, perm


@pytest.fixture(scope="module")
def per_gate(reference):
    """The per-gate GPU result, kept on the device for the blocked comparison."""
    st = DeviceState(N)
    brk.fill_hash(st.ptr, st.size, brk.C128, FILL_SEED, SCALE)
    assert st.size * 16 >= brk.get_tune("nt_min_bytes")  # the non-temporal kernels run
    note("30q: per-gate run")
    perm = brk.Circuit(N, _ops()).run(st.ptr)
    yield st, perm
    st.free()


def test_per_gate_vs_oracle(reference, per_gate):
    expected, perm_o = reference
    st, perm = per_gate
    assert perm == perm_o
    err = 0.0
    for start in range(0, st.size, CHUNK):
        got = _download_range(st, start, CHUNK)
        err = max(err, float(np.max(np.abs(got - expected[start:start + CHUNK]))))
    note(f"30q: per-gate max-abs vs oracle {err:.3e}")
    assert err < 1e-12


def test_blocked_bit_identical_to_per_gate(per_gate):
    plain, perm = per_gate
    st = DeviceState(N)
    try:
        brk.fill_hash(st.ptr, st.size, brk.C128, FILL_SEED, SCALE)
        note("30q: blocked run")
        perm_b = brk.Circuit(N, _ops()).run_blocked(st.ptr)
        assert perm_b == perm
        for start in range(0, st.size, CHUNK):
            a = _download_range(plain, start, CHUNK)
            b = _download_range(st, start, CHUNK)
            assert np.array_equal(a.view(np.uint64), b.view(np.uint64)), start
    finally:
        st.free()


@pytest.mark.parametrize("width", [4, 5, 6])
def test_fused_c64_vs_oracle(reference, width):
    expected, perm_o = reference
    src = DeviceState(N)
    st = DeviceState(N, brk.C64)
    try:
        brk.fill_hash(src.ptr, src.size, brk.C128, FILL_SEED, SCALE)
        brk.convert(st.ptr, brk.C64, src.ptr, brk.C128, st.size)
        src.free()
        plan = brk.FusedPlan(brk.Circuit(N, _ops()), width)
        assert plan.final_permutation() == perm_o
        note(f"30q: fused width {width}, {plan.num_blocks} blocks")
        plan.apply(st.ptr)
        dot = 0j
        nr = ng = 0.0
        for start in range(0, st.size, CHUNK):
            g = _download_range(st, start, CHUNK).astype(np.complex128)
            r = expected[start:start + CHUNK]
            dot += np.vdot(r, g)
            nr += np.vdot(r, r).real
            ng += np.vdot(g, g).real
        infidelity = 1.0 - abs(dot) ** 2 / (nr * ng)
        err = 0.0
        for start in range(0, st.size, CHUNK):
            g = _download_range(st, start, CHUNK).astype(np.complex128) / np.sqrt(ng)
            err = max(err, float(np.max(np.abs(g - expected[start:start + CHUNK] / np.sqrt(nr)))))
        note(f"30q: fused width {width}: 1-F {infidelity:.3e}, max-abs {err:.3e}")
        assert infidelity <= 1e-6
        assert err <= 2e-6
    finally:
        src.free()
        st.free()
