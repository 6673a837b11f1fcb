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
            assert ops[k][0] != "LOCAL_SWAP" and phys[k] <= set(local)
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


def test_single_ops_and_local_swaps_are_plain():
    ops = [("H", (0,), (), ()), ("LOCAL_SWAP", (1, 9), (), ()), ("H", (1,), (), ())]
    plan = brk.BlockPlan(brk.Circuit(10, ops))
    assert plan.steps == [(0, 1, ()), (1, 1, ()), (2, 1, ())]
