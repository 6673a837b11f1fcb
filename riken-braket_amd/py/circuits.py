"""Seeded synthetic workloads (the circuits the benchmark and the parity tests run).

Every generator returns plain op tuples (name, targets, controls, params) on
logical qubits, the input of brk.Circuit, plus emit_qcx() to write the same
list in bra's quantum-assembler syntax (docs/bra.md:109-188), so the CPU and
GPU paths read identical gate lists.  Angles are drawn with numpy's PCG64.

Choices the workload description leaves open (fixed here and used everywhere):
  * random circuit: exactly 60 % single-qubit / 40 % two-qubit gates in a
    seed-shuffled order; two-qubit gates are CNOT or CZ with p = 1/2 on a
    uniformly random distinct pair (CNOT: first = control); the local swaps
    go to seeded positions, each between a uniformly random physical qubit of
    the high half [n/2, n) and one of the low half [0, n/2).
  * RZ(theta) = EZ(-theta/2), RY(theta) = EY(-theta/2) (ket's conventions).
  * Ising Hamiltonians are open chains: sum_{i<n-1} J_i Z_i Z_{i+1} + sum_i h_i X_i.
"""
from __future__ import annotations

import math

import numpy as np

TWO_PI = 2.0 * math.pi


def layered_rz_cnot(n=20, layers=10, seed=1):
    """H on every qubit of |0..0>, then `layers` x [RZ(theta) on each qubit;
    CNOT(i, i+1) for even i in even layers, odd i in odd layers]."""
    rng = np.random.default_rng(seed)
    ops = [("H", (q,), (), ()) for q in range(n)]
    for layer in range(layers):
        for q in range(n):
            ops.append(("RZ", (q,), (), (float(rng.uniform(0.0, TWO_PI)),)))
        for i in range(layer % 2, n - 1, 2):
            ops.append(("X", (i + 1,), (i,), ()))
    return ops


def random_circuit(n=32, num_gates=200, num_local_swaps=8, seed=2):
    """60/40 single/two-qubit random gates plus local swaps (see module doc)."""
    rng = np.random.default_rng(seed)
    n1 = int(round(0.6 * num_gates))
    kinds = np.array([1] * n1 + [2] * (num_gates - n1))
    rng.shuffle(kinds)
    gates = []
    for k in kinds:
        if k == 1:
            g = ("H", "X", "RZ", "U3")[int(rng.integers(4))]
            t = int(rng.integers(n))
            if g == "RZ":
                gates.append(("RZ", (t,), (), (float(rng.uniform(0.0, TWO_PI)),)))
            elif g == "U3":
                gates.append(("U3", (t,), (), tuple(float(x) for x in rng.uniform(0.0, TWO_PI, 3))))
            else:
                gates.append((g, (t,), (), ()))
        else:
            a, b = (int(x) for x in rng.choice(n, 2, replace=False))
            if rng.integers(2) == 0:
                gates.append(("X", (b,), (a,), ()))  # CNOT(control a, target b)
            else:
                gates.append(("Z", (a, b), (), ()))  # CZ
    total = num_gates + num_local_swaps
    swap_slots = set(int(x) for x in rng.choice(total, num_local_swaps, replace=False))
    ops, it = [], iter(gates)
    half = n // 2
    for slot in range(total):
        if slot in swap_slots:
            hi = int(rng.integers(half, n))
            lo = int(rng.integers(0, half))
            ops.append(("LOCAL_SWAP", (hi, lo), (), ()))
        else:
            ops.append(next(it))
    return ops


def hea_layers(n, layers, rng, low=-math.pi, high=math.pi):
    """Hardware-efficient ansatz: layers x [RY(theta), RZ(phi) on each qubit; CZ ladder]."""
    ops = []
    for _ in range(layers):
        for q in range(n):
            ops.append(("RY", (q,), (), (float(rng.uniform(low, high)),)))
            ops.append(("RZ", (q,), (), (float(rng.uniform(low, high)),)))
        for i in range(n - 1):
            ops.append(("Z", (i, i + 1), (), ()))
    return ops


def batch_ansatz(num_circuits=4096, min_qubits=10, max_qubits=20, layers=8, seed=4):
    """[(n_c, ops, J, h)] for the batching form."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(num_circuits):
        n = int(rng.integers(min_qubits, max_qubits + 1))
        ops = hea_layers(n, layers, rng)
        J = rng.uniform(-1, 1, n - 1)
        h = rng.uniform(-1, 1, n)
        out.append((n, ops, J, h))
    return out


def adjoint_ansatz(n=24, layers=12, seed=5):
    """(ops, params) for the differentiation form: parameters are the RY/RZ angles in op order."""
    rng = np.random.default_rng(seed)
    ops = hea_layers(n, layers, rng)
    params = [op[3][0] for op in ops if op[0] in ("RY", "RZ")]
    return ops, params


def tfim(n, J=1.0, h=0.7):
    return np.full(n - 1, J), np.full(n, h)


PARAMETRIC = ("RX", "RY", "RZ", "EX", "EY", "EZ", "U1")


def parameters(ops):
    return [op[3][0] for op in ops if op[0] in PARAMETRIC]


def with_parameters(ops, params):
    """The same circuit with the angles of its parametric ops replaced, in op order."""
    it = iter(params)
    return [(name, t, c, (float(next(it)),)) if name in PARAMETRIC else (name, t, c, p) for name, t, c, p in ops]


def logical_ops(ops):
    """Drops local swaps: the logical circuit a permuted run must agree with."""
    return [op for op in ops if op[0] != "LOCAL_SWAP"]


def inverse_ops(ops):
    """The adjoint circuit (reverse order, each op inverted).  Local swaps undo
    themselves when replayed in reverse order."""
    out = []
    for name, t, c, p in reversed(ops):
        if name in ("H", "X", "Y", "Z", "SWAP", "LOCAL_SWAP"):
            out.append((name, t, c, p))
        elif name in ("RX", "RY", "RZ", "EX", "EY", "EZ", "U1"):
            out.append((name, t, c, (-p[0],)))
        elif name == "U3":
            out.append((name, t, c, (-p[0], -p[2], -p[1])))
        elif name == "U2":  # U2(phi, lam)^dagger = U3(-pi/2, -lam, -phi)
            out.append(("U3", t, c, (-math.pi / 2, -p[1], -p[0])))
        else:
            raise ValueError(name)
    return out


def algorithmic_bytes(op, n, amp_bytes=16):
    """2 * sizeof(amp) * (number of amplitudes the op can change)."""
    name, t, c, _ = op
    N = 1 << n
    if name in ("H", "U2", "U3", "RX", "RY", "RZ", "EX", "EY", "EZ", "MATRIX", "Y"):
        touched = N >> len(c)
    elif name == "X":
        touched = N >> len(c)
    elif name in ("Z", "U1"):
        touched = N >> (len(t) + len(c))
    elif name in ("SWAP", "LOCAL_SWAP"):
        touched = N >> (1 + len(c))
    else:
        raise ValueError(name)
    return 2 * amp_bytes * touched


def op_kind(op):
    """Kernel family an op runs on: 'pair' (butterfly / swap) or 'diag'."""
    return "diag" if op[0] in ("Z", "U1", "RZ", "EZ") else "pair"


def emit_qcx(n, ops):
    """bra quantum-assembler text for an op list (local swaps are emitted as SWAP of the
    logical qubits they exchange, which bra would execute as a data swap)."""
    lines = [f"QUBITS {n}"]
    for name, t, c, p in ops:
        if name == "H":
            lines.append(f"H {t[0]}")
        elif name == "X" and not c:
            lines.append("X " + " ".join(map(str, t)))
        elif name == "X" and len(c) == 1 and len(t) == 1:
            lines.append(f"CNOT {c[0]} {t[0]}")
        elif name == "Z" and len(t) + len(c) == 2:
            q = list(t) + list(c)
            lines.append(f"CZ {q[0]} {q[1]}")
        elif name == "RZ":
            lines.append(f"EZ {t[0]} {repr(-0.5 * p[0])}")
        elif name == "RY":
            lines.append(f"EY {t[0]} {repr(-0.5 * p[0])}")
        elif name == "U3":
            lines.append(f"U3 {t[0]} {repr(p[0])} {repr(p[1])} {repr(p[2])}")
        else:
            raise ValueError(f"no qcx form for {name}")
    return "\n".join(lines) + "\n"
