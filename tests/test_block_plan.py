"""CPU checks of the gate-block planner (brk_block_plan_*; no device needed).

Invariants: the steps cover the ops in order; a block holds at least two ops
and at most 64; its local qubits are ascending, min(L, n) of them (L = 12 for
complex<double>, 13 for complex<float>; no blocks when that is below 9), include the qubits addressing one
128-byte line and every physical qubit its ops touch; local swaps are plain
steps; the final permutation is the one the ops' local swaps produce
(ket::mpi::qubit_permutation::permutate, qubit_permutation.hpp:242-252)."""
import pytest

import brk
from randops import random_ops


def simulate(ops, n, perm0=None):
    """physical qubits of every op, and the final permutation, by the circuit runner's rules."""
    perm = list(range(n)) if perm0 is None else list(perm0)
    inv = [0] * n
    for q, p in enumerate(perm):
        inv[p] = q
    phys = []
    for name, t, c, _ in ops:
        if name == "LOCAL_SWAP":
            phys.append(set(t))
            la, lb = inv[t[0]], inv[t[1]]
            perm[la], perm[lb] = perm[lb], perm[la]
            inv[perm[la]], inv[perm[lb]] = la, lb
        else:
            phys.append({perm[q] for q in t + c})
    return phys, perm


@pytest.mark.parametrize("n,dtype,seed,swaps", [(3, brk.C128, 1, 0), (12, brk.C128, 2, 0), (13, brk.C64, 3, 2),
                                                (20, brk.C128, 4, 3), (32, brk.C128, 5, 8), (34, brk.C64, 6, 4)])
def test_plan_invariants(n, dtype, seed, swaps):
    ops = random_ops(n, 300, seed, local_swaps=swaps)
    perm0 = list(range(n))[::-1]
    plan = brk.BlockPlan(brk.Circuit(n, ops), dtype, perm0)
    phys, final = simulate(ops, n, perm0)
    L = min(12 if dtype == brk.C128 else 13, n)
    line = set(range(min(3 if dtype == brk.C128 else 4, n)))
    nxt = 0
    blocks = 0
    for first, count, local in plan.steps:
        assert first == nxt and count >= 1
        nxt += count
        if not local:
            assert count == 1
            continue
        blocks += 1
        assert 2 <= count <= 64
        assert list(local) == sorted(local) and len(local) == L
        assert line <= set(local)
        for k in range(first, first + count):
            assert phys[k] <= set(local)  # a local swap joins as the SWAP of its physical qubits
    assert nxt == len(ops) and (blocks > 0) == (L >= 9)  # a chunk holds at least 2^9 amplitudes
    assert plan.final_permutation == final


def test_blocks_are_maximal_runs():
    # 20 single-qubit gates on qubits 3..11 fill one block's 12 local qubits (the line
    # qubits 0..2 plus 3..11), so the gate on qubit 19 starts the next block
    n = 20
    ops = [("H", (3 + k % 9,), (), ()) for k in range(20)] + [("H", (19,), (), ()), ("X", (19,), (18,), ())]
    plan = brk.BlockPlan(brk.Circuit(n, ops))
    assert [(f, k) for f, k, _ in plan.steps] == [(0, 20), (20, 2)]
    assert plan.steps[0][2] == tuple(range(12))
    assert {18, 19} <= set(plan.steps[1][2])


def test_local_swaps_join_blocks_and_single_ops_are_plain():
    # all 10 qubits fit one chunk: the swap of physical 1 and 9 runs inside the block, and the
    # H on logical 1 after it acts on physical 9
    ops = [("H", (0,), (), ()), ("LOCAL_SWAP", (1, 9), (), ()), ("H", (1,), (), ())]
    plan = brk.BlockPlan(brk.Circuit(10, ops))
    assert plan.steps == [(0, 3, tuple(range(10)))]
    assert plan.final_permutation == [0, 9, 2, 3, 4, 5, 6, 7, 8, 1]
    # the swap moves logical 1 to physical 18, so the H after it joins the gates on 18 and 19
    ops = [("H", (19,), (), ()), ("LOCAL_SWAP", (1, 18), (), ()), ("H", (1,), (), ())]
    plan = brk.BlockPlan(brk.Circuit(20, ops))
    assert plan.steps == [(0, 3, (0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 18, 19))]
    # a lone op, and a lone local swap, are plain steps
    assert brk.BlockPlan(brk.Circuit(20, [("LOCAL_SWAP", (1, 18), (), ())])).steps == [(0, 1, ())]
    assert brk.BlockPlan(brk.Circuit(20, [("H", (19,), (), ())])).steps == [(0, 1, ())]


def _shares(a, b):
    qa, qb = set(a[1]) | set(a[2]), set(b[1]) | set(b[2])
    return a[0] == "LOCAL_SWAP" or b[0] == "LOCAL_SWAP" or bool(qa & qb)


@pytest.mark.parametrize("n,seed,swaps,L,line", [(10, 1, 0, 11, 3), (20, 2, 3, 11, 3), (24, 3, 0, 12, 4),
                                                 (6, 4, 2, 5, 2)])
def test_block_order_keeps_dependences(n, seed, swaps, L, line):
    """brk_circuit_block_order is a permutation that keeps every pair of ops sharing a
    qubit (and every op relative to a local swap) in the given order."""
    ops = random_ops(n, 250, seed, local_swaps=swaps)
    order = brk.Circuit(n, ops).block_order(L, line)
    assert sorted(order) == list(range(len(ops)))
    pos = {i: k for k, i in enumerate(order)}
    for i in range(len(ops)):
        for j in range(i + 1, len(ops)):
            if _shares(ops[i], ops[j]):
                assert pos[i] < pos[j], (i, j)


def test_block_order_packs_layered_ansatz():
    """24 qubits, 12 layers of RY, RZ on every qubit and a ZZ ladder: the given order
    needs 60 blocks of 12 local qubits; the block order needs fewer than half."""
    import circuits as W
    ops, _ = W.adjoint_ansatz(24, 12, 5)
    c = brk.Circuit(24, ops)
    before = brk.BlockPlan(c)
    after = brk.BlockPlan(c.reordered(c.block_order(11, 3)))
    nb = lambda p: sum(1 for _, _, local in p.steps if local)
    assert nb(before) == 60
    assert nb(after) < 30


def test_block_order_is_the_same_operator(oracle):
    """The reordered circuit gives the same state on the CPU oracle (ops on disjoint
    qubits commute): within 1e-12 of the given order's."""
    import numpy as np
    import dense as D
    from oracle_circuits import run_ops
    n = 10
    ops = random_ops(n, 200, 7)
    order = brk.Circuit(n, ops).block_order(6, 2)
    assert order != list(range(len(ops)))
    s = D.random_state(n, np.random.default_rng(7))
    a, b = s.copy(), s.copy()
    run_ops(oracle, a, ops)
    run_ops(oracle, b, [ops[i] for i in order])
    assert np.max(np.abs(a - b)) < 1e-12
