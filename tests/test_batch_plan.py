"""CPU checks of the batching-form planner (no device needed: planning does not
touch the GPU).  Executing the planned stages in numpy -- every op of a stage
applied in the order listed, gates restricted to the stage's local qubits --
must reproduce the CPU oracle's state energy for the original circuit, so the
reordering (commutation), the per-qubit gate products and the measurement
placement are all pinned here; the GPU tests pin the kernel that runs them."""
import numpy as np
import pytest

import brk
import circuits as W
from oracle_circuits import run_ops


def apply_1q(s, n, q, m):
    v = s.reshape(1 << (n - q - 1), 2, 1 << q)
    return np.einsum("ij,ajb->aib", m, v).reshape(-1)


def execute_plan(batch, c, n, J, h):
    s = np.zeros(1 << n, dtype=np.complex128)
    s[0] = 1
    idx = np.arange(1 << n, dtype=np.uint64)
    energy = 0.0
    for st in range(batch.num_stages(c)):
        local, ops = batch.stage(c, st)
        L = min(n, 13)
        assert len(local) == L and local == sorted(local)
        assert set(range(min(n, 4))) <= set(local)
        lmask = sum(1 << q for q in local)
        for kind, mask, mat in ops:
            if kind == brk.BATCH_GATE:
                assert mask & ~lmask == 0, "gate on a non-local qubit"
                assert bin(mask).count("1") == 1
                s = apply_1q(s, n, mask.bit_length() - 1, mat)
            elif kind == brk.BATCH_SIGN:
                sel = (idx & np.uint64(mask)) == np.uint64(mask)
                s[sel] = -s[sel]
            elif kind == brk.BATCH_MEASURE_X:
                assert mask & ~lmask == 0, "X measured on a non-local qubit"
                q = mask.bit_length() - 1
                v = s.reshape(1 << (n - q - 1), 2, 1 << q)
                energy += 2 * h[q] * float(np.sum(v[:, 0, :].conj() * v[:, 1, :]).real)
            elif kind == brk.BATCH_MEASURE_ZZ:
                p = np.abs(s) ** 2
                for i in range(n - 1):
                    zz = 1 - 2 * (((idx >> np.uint64(i)) ^ (idx >> np.uint64(i + 1))) & np.uint64(1)).astype(np.float64)
                    energy += J[i] * float(np.dot(p, zz))
            else:
                raise AssertionError(kind)
    return s, energy


def oracle_energy(oracle, n, ops, J, h):
    s = np.zeros(1 << n, dtype=np.complex128)
    s[0] = 1
    run_ops(oracle, s, ops)
    return s, oracle.expectation_ising(s, J, h)


@pytest.mark.parametrize("n,layers", [(3, 2), (10, 8), (14, 8), (15, 3), (17, 4), (20, 2)])
def test_plan_reproduces_oracle_energy(oracle, n, layers):
    rng = np.random.default_rng(100 + n)
    ops = W.hea_layers(n, layers, rng)
    J, h = rng.uniform(-1, 1, n - 1), rng.uniform(-1, 1, n)
    b = brk.Batch([(n, ops, J, h)])
    s_ref, e_ref = oracle_energy(oracle, n, ops, J, h)
    s, e = execute_plan(b, 0, n, J, h)
    assert np.max(np.abs(s - s_ref)) < 1e-12
    assert abs(e - e_ref) < 1e-12
    stages = b.num_stages(0)
    if n <= 13:
        assert stages == 1
    else:
        assert stages <= layers + 3  # about one stage per layer once n > 13


def test_plan_mixed_ops(oracle):
    # H/X/U3/RX/Z/CZ/CCZ in an irregular order, Z-type ops on arbitrary qubit sets
    n = 16
    rng = np.random.default_rng(7)
    ops = []
    for _ in range(120):
        k = int(rng.integers(6))
        q = int(rng.integers(n))
        if k == 0:
            ops.append(("H", (q,), (), ()))
        elif k == 1:
            ops.append(("U3", (q,), (), tuple(float(x) for x in rng.uniform(-3, 3, 3))))
        elif k == 2:
            ops.append(("RX", (q,), (), (float(rng.uniform(-3, 3)),)))
        elif k == 3:
            a, c = (int(x) for x in rng.choice(n, 2, replace=False))
            ops.append(("Z", (a, c), (), ()))
        elif k == 4:
            a, c, d = (int(x) for x in rng.choice(n, 3, replace=False))
            ops.append(("Z", (a,), (c, d), ()))
        else:
            ops.append(("X", (q,), (), ()))
    J, h = rng.uniform(-1, 1, n - 1), rng.uniform(-1, 1, n)
    b = brk.Batch([(n, ops, J, h)])
    s_ref, e_ref = oracle_energy(oracle, n, ops, J, h)
    s, e = execute_plan(b, 0, n, J, h)
    assert np.max(np.abs(s - s_ref)) < 1e-12
    assert abs(e - e_ref) < 1e-12


def test_plan_merges_gates_per_qubit():
    # RY then RZ on each qubit of one layer: one 2x2 gate per qubit
    n = 6
    rng = np.random.default_rng(1)
    ops = W.hea_layers(n, 1, rng)
    b = brk.Batch([(n, ops, np.ones(n - 1), np.ones(n))])
    _, sops = b.stage(0, 0)
    kinds = [k for k, _, _ in sops]
    assert kinds.count(brk.BATCH_GATE) == n
    assert kinds.count(brk.BATCH_SIGN) == n - 1
    assert kinds.count(brk.BATCH_MEASURE_X) == n and kinds.count(brk.BATCH_MEASURE_ZZ) == 1


def test_zero_hamiltonian_terms_are_not_measured():
    b = brk.Batch([(5, [("H", (0,), (), ())], np.zeros(4), np.array([0, 1.0, 0, 0, 0]))])
    _, sops = b.stage(0, 0)
    assert [(k, m) for k, m, _ in sops if k >= brk.BATCH_MEASURE_X] == [(brk.BATCH_MEASURE_X, 2)]


@pytest.mark.parametrize("op", [("X", (1,), (0,), ()), ("SWAP", (0, 1), (), ()), ("LOCAL_SWAP", (3, 0), (), ()),
                                ("X", (0, 1), (), ())])
def test_unsupported_ops_are_rejected(op):
    with pytest.raises(brk.BrkError) as e:
        brk.Batch([(4, [op], None, None)])
    assert e.value.status == 8  # BRK_ERR_UNSUPPORTED


def test_qubit_limits():
    with pytest.raises(brk.BrkError):
        brk.Batch([(25, [], None, None)])
    b = brk.Batch([(24, [("H", (23,), (), ())], None, np.ones(24))])
    assert b.num_stages(0) >= 2
