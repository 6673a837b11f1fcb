"""Independent numpy model of the gate set, used to pin the oracle.

Gates are built from their closed-form matrices (docs/bra.md:109-188 and the
comments above each ket gate function) and applied by their definition over
basis indices (vectorised over all indices at once) -- a different
algorithm from the index-mask group loop of the reference and the oracle, so
a shared mistake in the bit manipulation cannot hide.
"""
from __future__ import annotations

import cmath
import math

import numpy as np

R2 = 1.0 / math.sqrt(2.0)

I2 = np.eye(2, dtype=np.complex128)
H = np.array([[R2, R2], [R2, -R2]], dtype=np.complex128)
X = np.array([[0, 1], [1, 0]], dtype=np.complex128)
Y = np.array([[0, -1j], [1j, 0]], dtype=np.complex128)
Z = np.array([[1, 0], [0, -1]], dtype=np.complex128)
SWAP = np.array([[1, 0, 0, 0], [0, 0, 1, 0], [0, 1, 0, 0], [0, 0, 0, 1]], dtype=np.complex128)


def U1(lam):
    return np.array([[1, 0], [0, cmath.exp(1j * lam)]], dtype=np.complex128)


def U2(phi, lam):
    # docs/bra.md "U2 t phi lambda"
    return R2 * np.array([[1, -cmath.exp(1j * lam)],
                          [cmath.exp(1j * phi), cmath.exp(1j * (phi + lam))]], dtype=np.complex128)


def U3(theta, phi, lam):
    # docs/bra.md "U3 t theta phi lambda"
    c, s = math.cos(theta / 2), math.sin(theta / 2)
    return np.array([[c, -cmath.exp(1j * lam) * s],
                     [cmath.exp(1j * phi) * s, cmath.exp(1j * (phi + lam)) * c]], dtype=np.complex128)


def kron_all(mats):
    """kron over a list where mats[j] acts on local bit j (mats[0] = least significant)."""
    out = np.array([[1.0 + 0j]])
    for m in mats:
        out = np.kron(m, out)
    return out


def pauli_exp(theta, pauli, m):
    """exp(i theta P^(x)m) = cos I + i sin P^(x)m  (docs/bra.md EX/EY/EZ)."""
    P = kron_all([pauli] * m)
    return math.cos(theta) * np.eye(2 ** m, dtype=np.complex128) + 1j * math.sin(theta) * P


def RZ(theta):
    return np.diag([cmath.exp(-0.5j * theta), cmath.exp(0.5j * theta)]).astype(np.complex128)


def RY(theta):
    c, s = math.cos(theta / 2), math.sin(theta / 2)
    return np.array([[c, -s], [s, c]], dtype=np.complex128)


def apply(state, U, targets, controls=()):
    """Apply U under controls; returns a new array.

    Definition: for every basis index i whose control bits are all 1, with
    local value v(i) = sum_j bit_{targets[j]}(i) << j,
        out[i] = sum_u U[v(i), u] * state[(i with the target bits set to u)].
    """
    N = state.size
    idx = np.arange(N, dtype=np.int64)
    cmask = 0
    for c in controls:
        cmask |= 1 << c
    tmask = 0
    for t in targets:
        tmask |= 1 << t
    on = (idx & cmask) == cmask
    v = np.zeros(N, dtype=np.int64)
    for j, t in enumerate(targets):
        v |= ((idx >> t) & 1) << j
    base = idx & ~tmask
    new = np.zeros(N, dtype=np.complex128)
    for u in range(2 ** len(targets)):
        src = base.copy()
        for j, t in enumerate(targets):
            if (u >> j) & 1:
                src |= 1 << t
        new += U[v, u] * state[src]
    out = state.astype(np.complex128).copy()
    out[on] = new[on]
    return out


def full_matrix(n, U, targets, controls=()):
    """Dense 2^n x 2^n matrix of a gate (columns = images of basis states)."""
    N = 2 ** n
    M = np.zeros((N, N), dtype=np.complex128)
    for j in range(N):
        e = np.zeros(N, dtype=np.complex128)
        e[j] = 1
        M[:, j] = apply(e, U, targets, controls)
    return M


def ket_fixture_state(n=4):
    """ket/test/runtime_gate_numerical.cpp:42-50."""
    i = np.arange(2 ** n)
    return (0.125 * (i + 1) - 1j * 0.0625 * ((3 * i + 1) % 7)).astype(np.complex128)


def observable_state1(n=4):
    """ket/test/runtime_observable_numerical.cpp:30-38."""
    i = np.arange(2 ** n)
    return (0.0625 * ((5 * i + 3) % 17) - 1j * 0.03125 * ((7 * i + 1) % 13)).astype(np.complex128)


def observable_state2(n=4):
    """ket/test/runtime_observable_numerical.cpp:40-48."""
    i = np.arange(2 ** n)
    return (-0.046875 * ((11 * i + 2) % 19) + 1j * 0.078125 * ((3 * i + 5) % 11)).astype(np.complex128)


def random_state(n, rng, normalize=True):
    v = rng.standard_normal(2 ** n) + 1j * rng.standard_normal(2 ** n)
    if normalize:
        v /= np.linalg.norm(v)
    return v.astype(np.complex128)


def splitmix_unit(seed, counters):
    """numpy twin of the device fill_hash generator: values in [-1, 1)."""
    c = np.asarray(counters, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + (c + np.uint64(1)) * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0) * 2.0 - 1.0


def hash_state_values(seed, indices, scale=1.0):
    idx = np.asarray(indices, dtype=np.uint64)
    re = scale * splitmix_unit(seed, np.uint64(2) * idx)
    im = scale * splitmix_unit(seed, np.uint64(2) * idx + np.uint64(1))
    return re + 1j * im
