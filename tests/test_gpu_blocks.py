"""GPU parity of gate blocks (brk_circuit_run_blocked / brk_block_plan_run).

Bar: bit-identical amplitudes to the per-gate path (brk_circuit_run) -- a
block applies each gate with the same IEEE operation sequence -- and, for
complex<double>, within 1e-12 of the CPU oracle.  Circuits mix every op kind
with controls, multi-target strings, local swaps and a non-identity starting
permutation."""
import sys

import numpy as np
import pytest

import brk
import dense as D
from devstate import DeviceState
from oracle_circuits import run_ops
from randops import random_ops

pytestmark = pytest.mark.gpu


def note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def run_both(n, ops, dtype, perm0=None, seed=0):
    rng = np.random.default_rng(seed)
    s = D.random_state(n, rng)
    if dtype == brk.C64:
        s = s.astype(np.complex64).astype(np.complex128)
    c = brk.Circuit(n, ops)
    a, b = DeviceState.from_host(s, dtype), DeviceState.from_host(s, dtype)
    pa = c.run(a.ptr, dtype, perm0)
    pb = c.run_blocked(b.ptr, dtype, perm0)
    out_a, out_b = a.download(), b.download()
    a.free()
    b.free()
    return s, out_a, out_b, pa, pb


def same_bits(x, y):
    w = np.uint64 if x.dtype == np.complex128 else np.uint32
    return np.array_equal(x.view(w), y.view(w))


@pytest.fixture(params=["temporal", "nontemporal"])
def access(request):
    """Default access policy, then the non-temporal instantiations large states take."""
    if request.param == "temporal":
        yield request.param
        return
    old = brk.get_tune("nt_min_bytes")
    brk.tune("nt_min_bytes", 0)
    try:
        yield request.param
    finally:
        brk.tune("nt_min_bytes", old)


@pytest.mark.parametrize("n,dtype,seed", [(3, brk.C128, 1), (9, brk.C128, 2), (12, brk.C128, 3), (13, brk.C128, 4),
                                          (16, brk.C128, 5), (20, brk.C128, 6), (12, brk.C64, 7), (14, brk.C64, 8),
                                          (20, brk.C64, 9)])
def test_blocked_equals_plain_bitwise(oracle, n, dtype, seed, access):
    note(f"gpu blocks n={n} dtype={dtype} seed={seed} ({access})")
    ops = random_ops(n, 160, seed, local_swaps=3 if n >= 4 else 0)
    perm0 = list(np.random.default_rng(seed).permutation(n))
    s, plain, blocked, pa, pb = run_both(n, ops, dtype, perm0, seed)
    assert pa == pb
    assert same_bits(plain, blocked)
    if dtype == brk.C128:
        ref = s.copy()
        run_ops(oracle, ref, ops, perm0)
        assert np.max(np.abs(blocked - ref)) < 1e-12


def test_dense_low_qubit_blocks(oracle):
    """Many gates on the line qubits and their neighbours: long blocks, every pair mode."""
    n = 18
    note("gpu blocks: low-qubit circuit")
    ops = random_ops(n, 400, 11, qubit_pool=range(10))
    plan = brk.BlockPlan(brk.Circuit(n, ops))
    assert max(k for _, k, _ in plan.steps) == 64
    s, plain, blocked, _, _ = run_both(n, ops, brk.C128, None, 11)
    assert same_bits(plain, blocked)


def test_plan_reuse_and_large_state(oracle):
    """A 26-qubit random circuit (1 GiB), the plan run twice: each run equals the plain path."""
    n = 26
    note(f"gpu blocks: n={n} plan reuse")
    ops = random_ops(n, 60, 12, max_controls=1)
    c = brk.Circuit(n, ops)
    plan = brk.BlockPlan(c)
    st = DeviceState(n)
    ref = DeviceState(n)
    for dev in (st, ref):
        brk.fill_modular(dev.ptr, dev.size, brk.C128, (1e-4, 3e-4), (7, 3, 101, 11, 5, 97))
    plan.run(st.ptr)
    plan.run(st.ptr)
    c.run(ref.ptr)
    c.run(ref.ptr)
    brk.call("brk_stream_synchronize", None)
    x, y = st.download(), ref.download()
    assert same_bits(x, y)
    st.free()
    ref.free()


@pytest.mark.parametrize("n,dtype,seed", [(14, brk.C128, 21), (20, brk.C128, 22), (16, brk.C64, 23)])
def test_block_order_run_matches_plain(oracle, n, dtype, seed):
    """The circuit in block order (brk_circuit_block_order) run as gate blocks: fewer blocks,
    the same final permutation, amplitudes within rounding of the per-gate run of the given
    order (1e-12 complex128, 1e-5 complex64) and, for complex128, of the oracle."""
    note(f"gpu blocks: block order n={n} dtype={dtype} seed={seed}")
    ops = random_ops(n, 200, seed, local_swaps=4)
    c = brk.Circuit(n, ops)
    L = 12 if dtype == brk.C128 else 13
    ordered = c.reordered(c.block_order(L, 3 if dtype == brk.C128 else 4))
    nb = lambda p: sum(1 for _, _, local in p.steps if local)
    assert nb(brk.BlockPlan(ordered, dtype)) <= nb(brk.BlockPlan(c, dtype))
    rng = np.random.default_rng(seed)
    s = D.random_state(n, rng)
    if dtype == brk.C64:
        s = s.astype(np.complex64).astype(np.complex128)
    a, b = DeviceState.from_host(s, dtype), DeviceState.from_host(s, dtype)
    pa = c.run(a.ptr, dtype)
    pb = ordered.run_blocked(b.ptr, dtype)
    plain, blocked = a.download(), b.download()
    a.free()
    b.free()
    assert pa == pb
    tol = 1e-12 if dtype == brk.C128 else 1e-5
    assert np.max(np.abs(plain - blocked)) < tol
    if dtype == brk.C128:
        ref = s.copy()
        run_ops(oracle, ref, ops)
        assert np.max(np.abs(blocked - ref)) < 1e-12
