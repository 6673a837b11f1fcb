"""GPU checks of the batching form: energies of the persistent, graph-captured
batch kernel (complex64 states, FP64 energies) against the FP64 plain per-gate
path and the CPU oracle; determinism and graph replays."""
import sys

import numpy as np
import pytest

import brk
import circuits as W
from devstate import DeviceState
from oracle_circuits import run_ops

pytestmark = pytest.mark.gpu

# complex64 state, <= 8 layers: |E_c64 - E_c128| stays far below this
E_TOL = 2e-5


def note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def plain_energy(n, ops, J, h):
    st = DeviceState(n)
    brk.set_basis_state(st.ptr, st.size, brk.C128, 0)
    brk.Circuit(n, ops).run(st.ptr)
    e = brk.expectation_ising(st.ptr, st.size, brk.C128, J, h)
    st.free()
    return e


def test_small_batch_vs_plain_and_oracle(oracle):
    items = W.batch_ansatz(num_circuits=40, min_qubits=10, max_qubits=20, layers=8, seed=11)
    note(f"gpu batch: {len(items)} circuits, n = {sorted(set(n for n, _, _, _ in items))}")
    b = brk.Batch(items)
    e = b.run()
    for i, (n, ops, J, h) in enumerate(items):
        ref = plain_energy(n, ops, J, h)
        assert abs(e[i] - ref) < E_TOL, (i, n, e[i], ref)
        if n <= 12:
            s = np.zeros(1 << n, dtype=np.complex128)
            s[0] = 1
            run_ops(oracle, s, ops)
            assert abs(e[i] - oracle.expectation_ising(s, J, h)) < E_TOL


def test_mixed_ops_and_sizes():
    rng = np.random.default_rng(5)
    items = []
    for n in (2, 4, 5, 9, 14, 15, 16, 21):
        ops = []
        for _ in range(12 * n):
            k = int(rng.integers(5))
            q = int(rng.integers(n))
            if k == 0:
                ops.append(("H", (q,), (), ()))
            elif k == 1:
                ops.append(("U3", (q,), (), tuple(float(x) for x in rng.uniform(-3, 3, 3))))
            elif k == 2 and n >= 2:
                a, c = (int(x) for x in rng.choice(n, 2, replace=False))
                ops.append(("Z", (a, c), (), ()))
            elif k == 3 and n >= 3:
                a, c, d = (int(x) for x in rng.choice(n, 3, replace=False))
                ops.append(("Z", (a,), (c, d), ()))
            else:
                ops.append(("RY", (q,), (), (float(rng.uniform(-3, 3)),)))
        J = rng.uniform(-1, 1, max(0, n - 1))
        items.append((n, ops, J, rng.uniform(-1, 1, n)))
    note("gpu batch: mixed ops, n = 1..21")
    e = brk.Batch(items).run()
    for i, (n, ops, J, h) in enumerate(items):
        assert abs(e[i] - plain_energy(n, ops, J, h)) < E_TOL, (n, e[i])


def test_cz_terms_of_every_reach():
    """CZ between qubits at every distance d (sign terms x & (x >> d) & A), between and inside the
    gate groups, with U3 / H / RY around them so that every sign reaches the measured energy; one-
    and multi-stage circuits, complex and real gate forms, signs before and after the gates."""
    rng = np.random.default_rng(17)
    items = []
    for n in (3, 6, 11, 13, 14, 16, 19):
        ops = []
        for layer in range(6):
            for q in range(n):
                k = int(rng.integers(3))
                if k == 0:
                    ops.append(("U3", (q,), (), tuple(float(x) for x in rng.uniform(-3, 3, 3))))
                elif k == 1:
                    ops.append(("H", (q,), (), ()))
                else:
                    ops.append(("RY", (q,), (), (float(rng.uniform(-3, 3)),)))
            for _ in range(n):
                a, c = (int(x) for x in rng.choice(n, 2, replace=False))
                ops.append(("Z", (a,), (c,), ()))
        items.append((n, ops, rng.uniform(-1, 1, n - 1), rng.uniform(-1, 1, n)))
    note("gpu batch: CZ terms, n = 3..19")
    e = brk.Batch(items).run()
    for i, (n, ops, J, h) in enumerate(items):
        assert abs(e[i] - plain_energy(n, ops, J, h)) < E_TOL, (n, e[i])


def test_one_qubit_closed_form():
    # RY(t)|0> = cos(t/2)|0> + sin(t/2)|1>: <X> = sin t
    t, hx = 0.8, -0.6
    e = brk.Batch([(1, [("RY", (0,), (), (t,))], None, [hx])]).run()
    assert abs(e[0] - hx * np.sin(t)) < 1e-6


def test_deterministic_and_graph_replays():
    items = W.batch_ansatz(num_circuits=300, min_qubits=10, max_qubits=18, layers=4, seed=12)
    b = brk.Batch(items)
    note("gpu batch: replays")
    first = b.run()
    st = DeviceState(20)
    for r in range(3):
        # unrelated work on the device between replays
        brk.set_basis_state(st.ptr, st.size, brk.C128, r)
        brk.hadamard(st.ptr, st.size, brk.C128, 3)
        again, ms, t0, t1 = b.run_timed()
        assert np.array_equal(again, first)
        assert ms > 0
        assert np.all(t1 >= t0) and np.min(t0) == 0.0
        assert np.max(t1) <= ms * 1.05 + 0.05
    st.free()


def test_independent_batches_share_nothing():
    a_items = W.batch_ansatz(num_circuits=16, min_qubits=15, max_qubits=17, layers=3, seed=21)
    b_items = W.batch_ansatz(num_circuits=16, min_qubits=10, max_qubits=12, layers=3, seed=22)
    a, b = brk.Batch(a_items), brk.Batch(b_items)
    ea1, eb1 = a.run(), b.run()
    ea2 = a.run()
    assert np.array_equal(ea1, ea2)
    for i in (0, 5):
        n, ops, J, h = b_items[i]
        assert abs(eb1[i] - plain_energy(n, ops, J, h)) < E_TOL


@pytest.mark.slow
def test_full_workload_sampled():
    # the batching workload: 4096 circuits, 10..20 qubits, 8 layers
    items = W.batch_ansatz()
    note(f"gpu batch: {len(items)} circuits")
    e = brk.Batch(items).run()
    assert np.all(np.isfinite(e))
    rng = np.random.default_rng(3)
    for i in sorted(rng.choice(len(items), 24, replace=False)):
        n, ops, J, h = items[i]
        assert abs(e[i] - plain_energy(n, ops, J, h)) < E_TOL, (i, n)
