"""Circuit-level GPU parity: the benchmark workloads through the C ABI circuit runner.

* layered RZ/CNOT circuit, 20 qubits (the reference CPU case): GPU vs oracle, 1e-12
* random 60/40 circuit with local swaps at 20-22 qubits: GPU vs oracle, 1e-12,
  same final permutation
* permutation invariant physical[p] == logical[inverse_permutate_bits(perm, p)]
  (ket/test/mpi/runtime_gate_numerical.cpp:111-127): bit-exact
* 32 qubits (64 GiB): per-gate closed forms on sampled indices of a hashed
  state, and circuit followed by its inverse restores |0...0> within 1e-12
"""
import sys
import zlib

import numpy as np
import pytest

import brk
import circuits as W
import dense as D
import hip
from devstate import DeviceState
from oracle_circuits import run_ops

pytestmark = pytest.mark.gpu

TOL = 1e-12


def note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(params=["temporal", "nontemporal"])
def access(request):
    """Runs a test twice: with the default access policy (small states use ordinary
    loads and stores) and with the non-temporal kernel instantiations that large
    states take (tuning nt_min_bytes = 0)."""
    if request.param == "temporal":
        yield request.param
        return
    old = brk.get_tune("nt_min_bytes")
    brk.tune("nt_min_bytes", 0)
    try:
        yield request.param
    finally:
        brk.tune("nt_min_bytes", old)


def test_layered_rz_cnot_20q(oracle, access):
    note(f"gpu circuit: layered RZ/CNOT 20q ({access})")
    ops = W.layered_rz_cnot(20, 10, seed=1)
    assert len(ops) == 315
    s = np.zeros(1 << 20, dtype=np.complex128)
    s[0] = 1
    expected = s.copy()
    run_ops(oracle, expected, ops)
    st = DeviceState.from_host(s)
    W_c = brk.Circuit(20, ops)
    W_c.run(st.ptr)
    got = st.download()
    st.free()
    assert np.max(np.abs(got - expected)) < TOL
    assert abs(np.vdot(got, got).real - 1) < 1e-12


@pytest.mark.parametrize("n,seed", [(20, 2), (22, 3)])
def test_random_circuit_with_local_swaps(oracle, n, seed, access):
    note(f"gpu circuit: random n={n} seed={seed} ({access})")
    ops = W.random_circuit(n, 200, 8, seed=seed)
    rng = np.random.default_rng(seed)
    s = D.random_state(n, rng)
    expected = s.copy()
    perm_o, _ = run_ops(oracle, expected, ops)
    st = DeviceState.from_host(s)
    perm_g = brk.Circuit(n, ops).run(st.ptr)
    got = st.download()
    st.free()
    assert perm_g == perm_o
    assert np.max(np.abs(got - expected)) < TOL


def test_permutation_invariant_bit_exact(oracle):
    n = 20
    note("gpu circuit: permutation invariant")
    ops = W.random_circuit(n, 200, 8, seed=7)
    s = D.random_state(n, np.random.default_rng(7))
    a, b = DeviceState.from_host(s), DeviceState.from_host(s)
    perm = brk.Circuit(n, ops).run(a.ptr)
    brk.Circuit(n, W.logical_ops(ops)).run(b.ptr)
    phys, logical = a.download(), b.download()
    a.free()
    b.free()
    assert perm != list(range(n))
    inv = [0] * n
    for q, p in enumerate(perm):
        inv[p] = q
    idx = np.array([oracle.inverse_permutate_bits(perm, inv, p) for p in range(1 << n)])
    assert np.array_equal(phys, logical[idx])


def test_circuit_argument_errors():
    with pytest.raises(brk.BrkError):
        brk.Circuit(4, [("H", (4,), (), ())])
    with pytest.raises(brk.BrkError):
        brk.Circuit(4, [("U3", (0,), (), (1.0,))])
    with pytest.raises(brk.BrkError):
        brk.Circuit(4, [("X", (1,), (1,), ())])


# ---------------------------------------------------------------------------
# 32 qubits
# ---------------------------------------------------------------------------

N32 = 32


def _fits_32q():
    free, _ = hip.mem_info()
    return free > (68 << 30)


@pytest.fixture(scope="module")
def state32():
    if not _fits_32q():
        pytest.skip("needs 68 GiB of free device memory")
    st = DeviceState(N32)
    yield st
    st.free()


SEED = 1234
SCALE = 2.0 ** -16


def _sample_indices(rng, k=64):
    idx = set(int(x) for x in rng.integers(0, 1 << N32, k, dtype=np.uint64))
    idx |= {0, (1 << N32) - 1, 1 << 31, (1 << 31) - 1}
    return sorted(idx)


@pytest.mark.slow
@pytest.mark.parametrize("case", [
    ("H", 0), ("H", 31), ("H", 17), ("U3", 30), ("RZ", 31), ("X", 5), ("CNOT", (31, 0)), ("CNOT", (2, 29)),
    ("CZ", (30, 31)), ("CZ", (0, 31)), ("SWAP", (31, 3)), ("SWAP", (16, 30)),
])
def test_32q_gate_sampled_closed_form(state32, case):
    kind, q = case
    note(f"gpu 32q gate {case}")
    st = state32
    brk.fill_hash(st.ptr, st.size, brk.C128, SEED, SCALE)
    rng = np.random.default_rng(zlib.crc32(str(case).encode()))
    idx = np.array(_sample_indices(rng), dtype=np.uint64)
    val = lambda i: D.hash_state_values(SEED, i, SCALE)
    one = np.uint64(1)
    if kind == "H":
        brk.hadamard(st.ptr, st.size, brk.C128, q)
        m = np.uint64(1 << q)
        bit = (idx & m) != 0
        a0, a1 = val(idx & ~m), val(idx | m)
        r = 1 / np.sqrt(2)
        exp = np.where(bit, (a0 - a1) * r, (a0 + a1) * r)
    elif kind == "U3":
        th, ph, la = 0.7, 1.9, -2.2
        brk.phase_shift3(st.ptr, st.size, brk.C128, th, ph, la, q)
        U = D.U3(th, ph, la)
        m = np.uint64(1 << q)
        bit = ((idx & m) != 0).astype(int)
        a0, a1 = val(idx & ~m), val(idx | m)
        exp = U[bit, 0] * a0 + U[bit, 1] * a1
    elif kind == "RZ":
        th = 0.9
        brk.exponential_pauli_z(st.ptr, st.size, brk.C128, -th / 2, [q])
        bit = ((idx >> np.uint64(q)) & one).astype(bool)
        exp = val(idx) * np.where(bit, np.exp(0.5j * th), np.exp(-0.5j * th))
    elif kind == "X":
        brk.pauli_x(st.ptr, st.size, brk.C128, [q])
        exp = val(idx ^ np.uint64(1 << q))
    elif kind == "CNOT":
        c, t = q
        brk.controlled_not(st.ptr, st.size, brk.C128, t, [c])
        on = ((idx >> np.uint64(c)) & one).astype(bool)
        exp = np.where(on, val(idx ^ np.uint64(1 << t)), val(idx))
    elif kind == "CZ":
        a, b = q
        brk.pauli_z(st.ptr, st.size, brk.C128, [a, b])
        on = (((idx >> np.uint64(a)) & one) & ((idx >> np.uint64(b)) & one)).astype(bool)
        exp = np.where(on, -val(idx), val(idx))
    elif kind == "SWAP":
        a, b = q
        brk.swap_permutated_local_qubits(st.ptr, st.size, brk.C128, a, b)
        ba = (idx >> np.uint64(a)) & one
        bb = (idx >> np.uint64(b)) & one
        differ = ba != bb
        exp = np.where(differ, val(idx ^ np.uint64((1 << a) | (1 << b))), val(idx))
    got = st.download_indices(idx)
    assert np.max(np.abs(got - exp)) < TOL * SCALE * 4


@pytest.mark.slow
def test_32q_random_circuit_then_inverse(state32):
    note("gpu 32q circuit + inverse")
    st = state32
    ops = W.random_circuit(N32, 200, 8, seed=2)
    brk.set_basis_state(st.ptr, st.size, brk.C128, 0)
    perm = brk.Circuit(N32, ops).run(st.ptr)
    norm = brk.norm2(st.ptr, st.size, brk.C128)
    assert abs(norm - 1) < 1e-12
    back = brk.Circuit(N32, W.inverse_ops(ops)).run(st.ptr, permutation=perm)
    assert back == list(range(N32))
    a = st.download_indices([0, 1, 1 << 31, (1 << 32) - 1])
    assert abs(a[0] - 1) < TOL
    assert np.max(np.abs(a[1:])) < TOL
