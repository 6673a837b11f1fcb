"""Gate cases shared by the oracle tests (CPU) and the GPU parity tests.

Each case names one gate application and knows how to run it three ways:
  dense()          -> (U, targets, controls) for tests/dense.py (closed form)
  oracle(o, s)     -> apply with the CPU oracle to numpy state s
  gpu(ptr, size, dtype, stream) -> apply with the HIP engine through the C ABI

The first block mirrors the cases of the reference's
ket/test/runtime_gate_numerical.cpp:97-257 (4 qubits, fixed angles 0.25 / 0.375);
the generator `random_cases` adds every gate with random qubits/controls.
"""
from __future__ import annotations

import cmath
import math
import random
from dataclasses import dataclass, field

import numpy as np

import dense as D


@dataclass
class Case:
    name: str
    kind: str
    targets: list
    controls: list = field(default_factory=list)
    params: tuple = ()

    def __str__(self):
        return f"{self.name}:{self.kind}(t={self.targets},c={self.controls},p={self.params})"

    # -- closed form -------------------------------------------------------
    def dense(self):
        k, t, c, p = self.kind, self.targets, self.controls, self.params
        m = len(t)
        if k == "H":
            return D.H, t, c
        if k == "X":
            return D.kron_all([D.X] * m), t, c
        if k == "Y":
            return D.kron_all([D.Y] * m), t, c
        if k == "Z":  # Z on "control" qubits: -1 where all bits are 1
            q = t + c
            return np.diag([1.0] * (2 ** len(q) - 1) + [-1.0]).astype(np.complex128), q, []
        if k == "ZN":  # Z on every target under the controls
            return D.kron_all([D.Z] * m), t, c
        if k in ("SZN", "SZN+"):  # sqrt(Z) = diag(1, i) (adjoint diag(1, -i)) on every target
            sz = np.diag([1.0, 1j if k == "SZN" else -1j]).astype(np.complex128)
            return D.kron_all([sz] * m), t, c
        if k == "U1coeff":
            q = t + c
            return np.diag([1.0] * (2 ** len(q) - 1) + [p[0]]).astype(np.complex128), q, []
        if k == "U1":
            q = t + c
            return np.diag([1.0] * (2 ** len(q) - 1) + [cmath.exp(1j * p[0])]).astype(np.complex128), q, []
        if k == "U2":
            return D.U2(*p), t, c
        if k == "U3":
            return D.U3(*p), t, c
        if k == "EX":
            return D.pauli_exp(p[0], D.X, m), t, c
        if k == "EY":
            return D.pauli_exp(p[0], D.Y, m), t, c
        if k == "EZ":
            return D.pauli_exp(p[0], D.Z, m), t, c
        if k == "SWAP":
            return D.SWAP, t, c
        if k == "ESWAP":  # exp(i theta SWAP) = cos theta I + i sin theta SWAP
            return math.cos(p[0]) * np.eye(4) + 1j * math.sin(p[0]) * D.SWAP, t, c
        if k == "M1":
            return np.array(p, dtype=np.complex128).reshape(2, 2), t, c
        raise ValueError(k)

    # -- oracle ------------------------------------------------------------
    def oracle(self, o, s):
        k, t, c, p = self.kind, self.targets, self.controls, self.params
        if k == "H":
            o.hadamard(s, t[0], c)
        elif k == "X":
            o.pauli_x(s, t, c)
        elif k == "Y":
            o.pauli_y(s, t, c)
        elif k == "Z":
            o.pauli_z(s, t + c)
        elif k == "ZN":
            o.pauli_zn(s, t, c)
        elif k in ("SZN", "SZN+"):
            o.sqrt_pauli_zn(s, t, c, adjoint=k == "SZN+")
        elif k == "U1coeff":
            o.phase_shift_coeff(s, p[0], t + c)
        elif k == "U1":
            o.phase_shift(s, p[0], t + c)
        elif k == "U2":
            o.phase_shift2(s, p[0], p[1], t[0], c)
        elif k == "U3":
            o.phase_shift3(s, p[0], p[1], p[2], t[0], c)
        elif k == "EX":
            o.exponential_pauli_x(s, p[0], t, c)
        elif k == "EY":
            o.exponential_pauli_y(s, p[0], t, c)
        elif k == "EZ":
            o.exponential_pauli_z(s, p[0], t, c)
        elif k == "SWAP":
            o.swap(s, t[0], t[1], c)
        elif k == "ESWAP":
            o.exponential_swap(s, p[0], t[0], t[1], c)
        elif k == "M1":
            o.gate_1q(s, t[0], p, c)
        else:
            raise ValueError(k)

    # -- HIP engine --------------------------------------------------------
    def gpu(self, ptr, size, dtype, stream=0):
        import brk
        k, t, c, p = self.kind, self.targets, self.controls, self.params
        if k == "H":
            brk.hadamard(ptr, size, dtype, t[0], c, stream)
        elif k == "X":
            if len(t) == 1 and c:
                brk.controlled_not(ptr, size, dtype, t[0], c, stream)
            else:
                brk.pauli_x(ptr, size, dtype, t, c, stream)
        elif k == "Y":
            brk.pauli_y(ptr, size, dtype, t, c, stream)
        elif k == "Z":
            brk.pauli_z(ptr, size, dtype, t + c, stream)
        elif k == "ZN":
            brk.pauli_zn(ptr, size, dtype, t, c, stream)
        elif k == "SZN":
            brk.sqrt_pauli_zn(ptr, size, dtype, t, c, stream)
        elif k == "SZN+":
            brk.adj_sqrt_pauli_zn(ptr, size, dtype, t, c, stream)
        elif k == "U1coeff":
            brk.phase_shift_coeff(ptr, size, dtype, p[0], t + c, stream)
        elif k == "U1":
            brk.phase_shift(ptr, size, dtype, p[0], t + c, stream)
        elif k == "U2":
            brk.phase_shift2(ptr, size, dtype, p[0], p[1], t[0], c, stream)
        elif k == "U3":
            brk.phase_shift3(ptr, size, dtype, p[0], p[1], p[2], t[0], c, stream)
        elif k == "EX":
            brk.exponential_pauli_x(ptr, size, dtype, p[0], t, c, stream)
        elif k == "EY":
            brk.exponential_pauli_y(ptr, size, dtype, p[0], t, c, stream)
        elif k == "EZ":
            brk.exponential_pauli_z(ptr, size, dtype, p[0], t, c, stream)
        elif k == "SWAP":
            brk.swap(ptr, size, dtype, t[0], t[1], c, stream)
        elif k == "ESWAP":
            brk.exponential_swap(ptr, size, dtype, p[0], t[0], t[1], c, stream)
        elif k == "M1":
            brk.gate_1q(ptr, size, dtype, t[0], p, c, stream)
        else:
            raise ValueError(k)


R2 = 1.0 / math.sqrt(2.0)

# ket/test/runtime_gate_numerical.cpp:97-257 (same gates, qubits, controls and angles;
# controlled_v (controlled_v.hpp:30-90: V(theta) = [[1+e, 1-e], [1-e, 1+e]] / 2 with
# e = e^{i theta}) and x/y_rotation_half_pi are given as explicit matrices through
# gate_1q, sqrt_pauli_z as the phase coefficient i).
_E025 = cmath.exp(0.25j)
REFERENCE_CASES = [
    Case("ref:hadamard", "H", [3], [0]),
    Case("ref:pauli_x", "X", [0, 3], [1]),
    Case("ref:pauli_y", "Y", [0, 3], [1]),
    Case("ref:pauli_z", "Z", [3], [0]),
    Case("ref:phase_shift_coeff", "U1coeff", [3], [0], (cmath.exp(0.375j),)),
    Case("ref:sqrt_pauli_z", "U1coeff", [3], [0], (1j,)),
    Case("ref:swap", "SWAP", [0, 3], [1]),
    Case("ref:x_rotation_half_pi", "M1", [3], [0], (R2, 1j * R2, 1j * R2, R2)),
    Case("ref:y_rotation_half_pi", "M1", [3], [0], (R2, R2, -R2, R2)),
    Case("ref:exponential_pauli_x", "EX", [0, 3], [1], (0.25,)),
    Case("ref:exponential_pauli_y", "EY", [0, 3], [1], (0.25,)),
    Case("ref:exponential_pauli_z", "EZ", [0, 3], [1], (0.25,)),
    Case("ref:controlled_v", "M1", [3], [0, 1], ((1 + _E025) / 2, (1 - _E025) / 2, (1 - _E025) / 2, (1 + _E025) / 2)),
    Case("ref:exponential_swap", "ESWAP", [0, 3], [1], (0.25,)),
]

KINDS_1T = ["H", "U2", "U3", "M1"]  # single target with controls
KINDS_MT = ["X", "Y", "EX", "EY", "EZ", "ZN", "SZN", "SZN+"]  # any number of targets with controls
KINDS_Q = ["Z", "U1", "U1coeff"]  # "control-type" qubits only


def _params(kind, rng):
    a = lambda: rng.uniform(0, 2 * math.pi)
    if kind == "U2":
        return (a(), a())
    if kind == "U3":
        return (a(), a(), a())
    if kind in ("EX", "EY", "EZ", "U1", "ESWAP"):
        return (a(),)
    if kind == "U1coeff":
        return (cmath.exp(1j * a()),)
    if kind == "M1":
        m = np.array([complex(rng.gauss(0, 1), rng.gauss(0, 1)) for _ in range(4)])
        return tuple(m)
    return ()


def random_cases(n, count, seed, max_targets=3, max_controls=2):
    rng = random.Random(seed)
    kinds = KINDS_1T + KINDS_MT + KINDS_Q + (["SWAP", "ESWAP"] if n >= 2 else [])
    out = []
    for i in range(count):
        kind = kinds[i % len(kinds)]
        if kind in KINDS_1T:
            nt = 1
        elif kind in ("SWAP", "ESWAP"):
            nt = 2
        elif kind in KINDS_Q:
            nt = rng.randint(1, min(3, n))
        else:
            nt = rng.randint(1, min(max_targets, n))
        nc = rng.randint(0, min(max_controls, n - nt))
        qs = rng.sample(range(n), nt + nc)
        out.append(Case(f"rand{i}", kind, qs[:nt], qs[nt:], _params(kind, rng)))
    return out
