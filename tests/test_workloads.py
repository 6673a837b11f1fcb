"""CPU checks of the workload generators and the oracle circuit runner."""
import numpy as np

import circuits as W
import dense as D
from oracle_circuits import run_ops


def test_layered_rz_cnot_shape():
    ops = W.layered_rz_cnot(20, 10, seed=1)
    assert len(ops) == 20 + 200 + 95
    assert ops == W.layered_rz_cnot(20, 10, seed=1)


def test_random_circuit_composition():
    ops = W.random_circuit(32, 200, 8, seed=2)
    assert len(ops) == 208
    gates = W.logical_ops(ops)
    one = [g for g in gates if len(g[1]) + len(g[2]) == 1]
    two = [g for g in gates if len(g[1]) + len(g[2]) == 2]
    assert len(one) == 120 and len(two) == 80
    swaps = [op for op in ops if op[0] == "LOCAL_SWAP"]
    assert len(swaps) == 8
    assert all(h >= 16 and l < 16 for _, (h, l), _, _ in swaps)
    total = sum(W.algorithmic_bytes(op, 32) for op in ops)
    assert 19e12 < total < 23e12  # about 21 TB of algorithmic traffic


def test_oracle_runner_with_local_swaps_matches_dense(oracle):
    # physical result of a permuted run == logical dense result re-indexed by the permutation
    n = 6
    ops = W.random_circuit(n, 40, 4, seed=11)
    s = D.random_state(n, np.random.default_rng(1))
    phys = s.copy()
    perm, _ = run_ops(oracle, phys, ops)
    logical = s.copy()
    for name, t, c, p in W.logical_ops(ops):
        if name == "X":
            logical = D.apply(logical, D.X, list(t), list(c))
        elif name == "Z":
            logical = D.apply(logical, np.diag([1, 1, 1, -1]).astype(complex), list(t))
        elif name == "H":
            logical = D.apply(logical, D.H, list(t))
        elif name == "RZ":
            logical = D.apply(logical, D.RZ(p[0]), list(t))
        elif name == "U3":
            logical = D.apply(logical, D.U3(*p), list(t))
    inv = [0] * n
    for q, ph in enumerate(perm):
        inv[ph] = q
    for ph_idx in range(1 << n):
        lg = sum(((ph_idx >> perm[q]) & 1) << q for q in range(n))
        assert abs(phys[ph_idx] - logical[lg]) < 1e-12


def test_inverse_ops_restore_state(oracle):
    n = 7
    ops = W.random_circuit(n, 60, 4, seed=5) + W.hea_layers(n, 2, np.random.default_rng(3))
    s = D.random_state(n, np.random.default_rng(2))
    t = s.copy()
    perm, _ = run_ops(oracle, t, ops)
    perm2, _ = run_ops(oracle, t, W.inverse_ops(ops), perm)
    assert perm2 == list(range(n))
    assert np.max(np.abs(t - s)) < 1e-12


def test_qcx_emission():
    text = W.emit_qcx(3, [("H", (0,), (), ()), ("X", (2,), (0,), ()), ("Z", (1, 2), (), ()),
                          ("RZ", (1,), (), (0.5,)), ("U3", (2,), (), (0.1, 0.2, 0.3))])
    assert text.splitlines() == ["QUBITS 3", "H 0", "CNOT 0 2", "CZ 1 2", "EZ 1 -0.25", "U3 2 0.1 0.2 0.3"]
