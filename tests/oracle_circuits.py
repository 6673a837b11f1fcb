"""Runs op lists (riken-braket_amd/py/circuits.py format) through the CPU oracle.

Test infrastructure (checker / CPU baseline).  Same semantics as the engine's
circuit runner: ops address logical qubits mapped through the permutation;
LOCAL_SWAP swaps physical data and calls permutate on the logical qubits
(ket/include/ket/mpi/utility/detail/make_local_swap_qubit.hpp:62-66).
"""
from __future__ import annotations

import time


def apply_op(o, s, op, perm):
    name, t, c, p = op
    if name == "LOCAL_SWAP":
        o.swap_permutated_local_qubits(s, t[0], t[1])
        return
    pt = [perm[q] for q in t]
    pc = [perm[q] for q in c]
    if name == "H":
        o.hadamard(s, pt[0], pc)
    elif name == "X":
        o.pauli_x(s, pt, pc)
    elif name == "Y":
        o.pauli_y(s, pt, pc)
    elif name == "Z":
        o.pauli_z(s, pt + pc)
    elif name == "U1":
        o.phase_shift(s, p[0], pt + pc)
    elif name == "PHASE":
        o.phase_shift_coeff(s, complex(p[0], p[1]), pt + pc)
    elif name == "U2":
        o.phase_shift2(s, p[0], p[1], pt[0], pc)
    elif name == "U3":
        o.phase_shift3(s, p[0], p[1], p[2], pt[0], pc)
    elif name == "EX":
        o.exponential_pauli_x(s, p[0], pt, pc)
    elif name == "EY":
        o.exponential_pauli_y(s, p[0], pt, pc)
    elif name == "EZ":
        o.exponential_pauli_z(s, p[0], pt, pc)
    elif name == "RX":
        o.exponential_pauli_x(s, -0.5 * p[0], pt, pc)
    elif name == "RY":
        o.exponential_pauli_y(s, -0.5 * p[0], pt, pc)
    elif name == "RZ":
        o.exponential_pauli_z(s, -0.5 * p[0], pt, pc)
    elif name == "SWAP":
        o.swap(s, pt[0], pt[1], pc)
    elif name == "MATRIX":
        o.gate_1q(s, pt[0], [complex(p[2 * i], p[2 * i + 1]) for i in range(4)], pc)
    elif name == "ESWAP":
        o.exponential_swap(s, p[0], pt[0], pt[1], pc)
    else:
        raise ValueError(name)


def run_ops(o, s, ops, perm=None, time_budget=None):
    """Applies ops in place; returns (final permutation, number of ops applied).
    With time_budget (seconds) stops after the op that crosses it."""
    n = s.size.bit_length() - 1
    perm = list(range(n)) if perm is None else list(perm)
    inv = [0] * n
    for q, ph in enumerate(perm):
        inv[ph] = q
    t0 = time.perf_counter()
    done = 0
    for op in ops:
        apply_op(o, s, op, perm)
        if op[0] == "LOCAL_SWAP":
            o.permutate(perm, inv, inv[op[1][0]], inv[op[1][1]])
        done += 1
        if time_budget is not None and time.perf_counter() - t0 > time_budget:
            break
    return perm, done
