"""Random op lists over every op kind of the circuit API (test infrastructure).

Ops are (name, targets, controls, params) tuples of riken-braket_amd/py/brk.py.
`local_swaps` > 0 inserts LOCAL_SWAP ops of physical qubits."""
from __future__ import annotations

import math
import random

ONE_TARGET = ("H", "U2", "U3", "MATRIX")
MULTI_TARGET = ("X", "Y", "EX", "EY", "EZ", "RX", "RY", "RZ")
QUBITS_ONLY = ("Z", "U1", "PHASE")
PAIR = ("SWAP", "ESWAP")
KINDS = ONE_TARGET + MULTI_TARGET + QUBITS_ONLY + PAIR


def _params(rng, name):
    a = lambda: rng.uniform(-math.pi, math.pi)
    if name in ("EX", "EY", "EZ", "RX", "RY", "RZ", "U1", "ESWAP"):
        return (a(),)
    if name == "U2":
        return (a(), a())
    if name == "U3":
        return (a(), a(), a())
    if name == "MATRIX":
        return tuple(rng.gauss(0, 0.7) for _ in range(8))
    if name == "PHASE":  # a unit coefficient, exactly i or -i half of the time (S, S+)
        t = rng.choice([math.pi / 2, -math.pi / 2, a()])
        return (0.0, math.copysign(1.0, t)) if abs(abs(t) - math.pi / 2) < 1e-15 else (math.cos(t), math.sin(t))
    return ()


def random_ops(n, count, seed, max_targets=3, max_controls=2, local_swaps=0, qubit_pool=None):
    """qubit_pool: restrict operated qubits to these (default: all n)."""
    rng = random.Random(seed)
    pool = list(range(n)) if qubit_pool is None else list(qubit_pool)
    ops = []
    for i in range(count):
        name = KINDS[rng.randrange(len(KINDS))]
        if name in ONE_TARGET:
            nt = 1
        elif name in PAIR:
            nt = 2
        elif name in QUBITS_ONLY:
            nt = rng.randint(1, min(3, len(pool)))
        else:
            nt = rng.randint(1, min(max_targets, len(pool)))
        if nt > len(pool):
            continue
        nc = rng.randint(0, min(max_controls, len(pool) - nt))
        qs = rng.sample(pool, nt + nc)
        ops.append((name, tuple(qs[:nt]), tuple(qs[nt:]), _params(rng, name)))
    for _ in range(local_swaps):
        a, b = rng.sample(range(n), 2)
        ops.insert(rng.randrange(len(ops) + 1), ("LOCAL_SWAP", (a, b), (), ()))
    return ops
