"""CPU checks of the fusion planner (precision form, host side).

The planned blocks, applied in order with their FP64 matrices by the
independent numpy model, must reproduce the oracle's per-gate result of the
circuit (1e-12): this pins block membership (commutation-respecting reorder),
the in-order product, identity widening and the local-swap -> SWAP mapping.
"""
import numpy as np
import pytest

import brk
import circuits as W
import dense as D
from oracle_circuits import run_ops


def _blocks_result(plan, s):
    out = s.copy()
    for i in range(plan.num_blocks):
        q, _, m = plan.block(i)
        out = D.apply(out, m, q)
    return out


@pytest.mark.parametrize("width", [4, 5, 6])
@pytest.mark.parametrize("seed", [1, 2])
def test_plan_reproduces_circuit(oracle, width, seed):
    n = 12
    ops = W.random_circuit(n, 80, 4, seed=seed) + W.hea_layers(n, 1, np.random.default_rng(seed))
    c = brk.Circuit(n, ops)
    plan = brk.FusedPlan(c, width)
    s = D.random_state(n, np.random.default_rng(seed))
    expected = s.copy()
    perm, _ = run_ops(oracle, expected, ops)
    assert plan.final_permutation() == perm
    got = _blocks_result(plan, s)
    assert np.max(np.abs(got - expected)) < 1e-12
    assert plan.num_blocks < len(ops)
    total = sum(plan.block(i, with_matrix=False)[1] for i in range(plan.num_blocks))
    assert total == len(ops)


def test_block_matrices_unitary():
    n = 12
    ops = W.random_circuit(n, 60, 2, seed=9)
    plan = brk.FusedPlan(brk.Circuit(n, ops), 5)
    for i in range(plan.num_blocks):
        q, g, m = plan.block(i)
        assert q == sorted(q) and len(q) == 5 and g >= 1
        assert np.max(np.abs(m @ m.conj().T - np.eye(32))) < 1e-12


def test_fusion_counts_32q():
    # the benchmark circuit: far fewer blocks than ops
    ops = W.random_circuit(32, 200, 8, seed=2)
    c = brk.Circuit(32, ops)
    counts = {w: brk.FusedPlan(c, w).num_blocks for w in (4, 5, 6)}
    assert counts[4] > counts[5] >= counts[6]
    assert counts[6] < 40


def test_plan_errors():
    c = brk.Circuit(8, [("H", (0,), (), ())])
    with pytest.raises(brk.BrkError):
        brk.FusedPlan(c, 4)  # needs n >= width + 5
    c = brk.Circuit(12, [("H", (0,), (), ())])
    with pytest.raises(brk.BrkError):
        brk.FusedPlan(c, 7)


def test_plans_are_reproducible():
    """The planner keeps the fewest-block candidate among orders drawn with fixed seeds: two
    plans of one circuit are identical (qubits, gate counts, matrices)."""
    ops = W.random_circuit(32, 200, 8, seed=2)
    c = brk.Circuit(32, ops)
    for w in (4, 5, 6):
        a, b = brk.FusedPlan(c, w), brk.FusedPlan(c, w)
        assert a.num_blocks == b.num_blocks
        for i in range(a.num_blocks):
            qa, ga, ma = a.block(i)
            qb, gb, mb = b.block(i)
            assert qa == qb and ga == gb and np.array_equal(ma, mb)
    # the 32-qubit benchmark circuit: 28 blocks at width 5 (32 in its given order without reordering)
    assert brk.FusedPlan(c, 5).num_blocks <= 28
