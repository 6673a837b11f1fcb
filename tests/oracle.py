"""ctypes binding of the CPU oracle (oracle/libket_oracle.so).

Test infrastructure: used by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg only, as the checker.  States are numpy complex128 arrays
(the memory layout of std::complex<double>[]).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")

_u = C.c_uint
_u64 = C.c_uint64
_d = C.c_double
_p = C.c_void_p
_up = C.POINTER(C.c_uint)

_SIGS = {
    "or_set_num_threads": ([C.c_int], None),
    "or_get_num_threads": ([], C.c_int),
    "or_make_index_masks": ([_up, _u, C.POINTER(C.c_uint64)], None),
    "or_index_with_qubits": ([_u64, _u64, _up, _u, C.POINTER(C.c_uint64)], C.c_uint64),
    "or_hadamard": ([_p, _u64, _u, _up, _u], None),
    "or_pauli_x": ([_p, _u64, _up, _u, _up, _u], None),
    "or_pauli_y": ([_p, _u64, _up, _u, _up, _u], None),
    "or_pauli_z": ([_p, _u64, _up, _u], None),
    "or_pauli_zn": ([_p, _u64, _up, _u, _up, _u], None),
    "or_sqrt_pauli_zn": ([_p, _u64, _up, _u, _up, _u, C.c_int], None),
    "or_phase_shift_coeff": ([_p, _u64, C.POINTER(C.c_double), _up, _u], None),
    "or_phase_shift2": ([_p, _u64, _d, _d, _u, _up, _u], None),
    "or_phase_shift3": ([_p, _u64, _d, _d, _d, _u, _up, _u], None),
    "or_exponential_pauli_x": ([_p, _u64, _d, _up, _u, _up, _u], None),
    "or_exponential_pauli_y": ([_p, _u64, _d, _up, _u, _up, _u], None),
    "or_exponential_pauli_z": ([_p, _u64, _d, _up, _u, _up, _u], None),
    "or_swap": ([_p, _u64, _u, _u, _up, _u], None),
    "or_exponential_swap": ([_p, _u64, _d, _u, _u, _up, _u], None),
    "or_gate_1q": ([_p, _u64, _u, C.POINTER(C.c_double), _up, _u], None),
    "or_swap_permutated_local_qubits": ([_p, _u64, _u, _u], None),
    "or_permutate": ([_up, _up, _u, _u, _u], None),
    "or_permutate_bits": ([_up, _u, _u64], C.c_uint64),
    "or_inverse_permutate_bits": ([_up, _up, _u, _u64], C.c_uint64),
    "or_inner_product": ([_p, _p, _u64, C.POINTER(C.c_double)], None),
    "or_norm2": ([_p, _u64], C.c_double),
    "or_spin_expectation_value": ([_p, _u64, _u, C.POINTER(C.c_double)], None),
    "or_expectation_pauli_string": ([_p, _u64, C.c_char_p, _up, _u, C.POINTER(C.c_double)], None),
    "or_expectation_ising": ([_p, _u64, C.POINTER(C.c_double), C.POINTER(C.c_double)], C.c_double),
    "or_apply_ising": ([_p, _p, _u64, C.POINTER(C.c_double), C.POINTER(C.c_double)], None),
    "or_mt64_seed": ([_p, _u64], None),
    "or_mt64_next": ([_p], C.c_uint64),
    "or_mt64_canonical": ([_p], C.c_double),
    "or_zero_one_probabilities": ([_p, _u64, _u, C.POINTER(C.c_double)], None),
    "or_project": ([_p, _u64, _u, C.c_int, _d], None),
    "or_projective_measurement": ([_p, _u64, _u, _d], C.c_int),
    "or_generate_events": ([_p, _u64, C.POINTER(C.c_double), _u, C.POINTER(C.c_uint64)], None),
    "or_shor_box": ([_p, _u64, _u, _u64, _u64], None),
}


class _Mt64State(C.Structure):
    _fields_ = [("mt", C.c_uint64 * 312), ("index", C.c_uint)]

_libs = {}


def load(fast: bool = False):
    """fast=False: -O3 checker build; fast=True: the reference's -Ofast release flags."""
    key = "fast" if fast else "check"
    if key in _libs:
        return _libs[key]
    name = "libket_oracle_fast.so" if fast else "libket_oracle.so"
    path = os.path.join(ORACLE_DIR, name)
    if not os.path.exists(path):
        raise RuntimeError(f"{path} not built (make -C oracle)")
    lib = C.CDLL(path)
    for fn, (args, res) in _SIGS.items():
        f = getattr(lib, fn)
        f.argtypes = args
        f.restype = res
    _libs[key] = lib
    return lib


def _ua(values):
    values = list(values)
    if not values:
        return None, 0
    return (C.c_uint * len(values))(*values), len(values)


def _da(values):
    values = list(values)
    return (C.c_double * len(values))(*values)


def _ptr(a: np.ndarray):
    assert a.dtype == np.complex128 and a.flags.c_contiguous
    return a.ctypes.data


class Oracle:
    """Method-per-function wrapper over the oracle library, mutating numpy states in place."""

    def __init__(self, fast: bool = False):
        self.lib = load(fast)

    def set_num_threads(self, n):
        self.lib.or_set_num_threads(n)

    def num_threads(self):
        return self.lib.or_get_num_threads()

    def hadamard(self, s, target, controls=()):
        c, nc = _ua(controls)
        self.lib.or_hadamard(_ptr(s), s.size, target, c, nc)

    def pauli_x(self, s, targets, controls=()):
        t, nt = _ua(targets)
        c, nc = _ua(controls)
        self.lib.or_pauli_x(_ptr(s), s.size, t, nt, c, nc)

    def controlled_not(self, s, target, controls):
        self.pauli_x(s, [target], controls)

    def pauli_y(self, s, targets, controls=()):
        t, nt = _ua(targets)
        c, nc = _ua(controls)
        self.lib.or_pauli_y(_ptr(s), s.size, t, nt, c, nc)

    def pauli_z(self, s, qubits):
        q, nq = _ua(qubits)
        self.lib.or_pauli_z(_ptr(s), s.size, q, nq)

    def pauli_zn(self, s, targets, controls=()):
        t, nt = _ua(targets)
        c, nc = _ua(controls)
        self.lib.or_pauli_zn(_ptr(s), s.size, t, nt, c, nc)

    def sqrt_pauli_zn(self, s, targets, controls=(), adjoint=False):
        t, nt = _ua(targets)
        c, nc = _ua(controls)
        self.lib.or_sqrt_pauli_zn(_ptr(s), s.size, t, nt, c, nc, 1 if adjoint else 0)

    def phase_shift_coeff(self, s, coeff, qubits):
        q, nq = _ua(qubits)
        self.lib.or_phase_shift_coeff(_ptr(s), s.size, _da([coeff.real, coeff.imag]), q, nq)

    def phase_shift(self, s, phase, qubits):
        import cmath
        self.phase_shift_coeff(s, cmath.exp(1j * phase), qubits)

    def phase_shift2(self, s, p1, p2, target, controls=()):
        c, nc = _ua(controls)
        self.lib.or_phase_shift2(_ptr(s), s.size, p1, p2, target, c, nc)

    def phase_shift3(self, s, p1, p2, p3, target, controls=()):
        c, nc = _ua(controls)
        self.lib.or_phase_shift3(_ptr(s), s.size, p1, p2, p3, target, c, nc)

    def exponential_pauli_x(self, s, phase, targets, controls=()):
        t, nt = _ua(targets)
        c, nc = _ua(controls)
        self.lib.or_exponential_pauli_x(_ptr(s), s.size, phase, t, nt, c, nc)

    def exponential_pauli_y(self, s, phase, targets, controls=()):
        t, nt = _ua(targets)
        c, nc = _ua(controls)
        self.lib.or_exponential_pauli_y(_ptr(s), s.size, phase, t, nt, c, nc)

    def exponential_pauli_z(self, s, phase, targets, controls=()):
        t, nt = _ua(targets)
        c, nc = _ua(controls)
        self.lib.or_exponential_pauli_z(_ptr(s), s.size, phase, t, nt, c, nc)

    def swap(self, s, q1, q2, controls=()):
        c, nc = _ua(controls)
        self.lib.or_swap(_ptr(s), s.size, q1, q2, c, nc)

    def exponential_swap(self, s, phase, q1, q2, controls=()):
        c, nc = _ua(controls)
        self.lib.or_exponential_swap(_ptr(s), s.size, phase, q1, q2, c, nc)

    def gate_1q(self, s, target, matrix, controls=()):
        m = _da([v for z in matrix for v in (z.real, z.imag)])
        c, nc = _ua(controls)
        self.lib.or_gate_1q(_ptr(s), s.size, target, m, c, nc)

    def swap_permutated_local_qubits(self, s, p1, p2):
        self.lib.or_swap_permutated_local_qubits(_ptr(s), s.size, p1, p2)

    def permutate(self, perm, inverse, q1, q2):
        """perm/inverse: lists, updated in place."""
        n = len(perm)
        pa = (C.c_uint * n)(*perm)
        ia = (C.c_uint * n)(*inverse)
        self.lib.or_permutate(pa, ia, n, q1, q2)
        perm[:] = list(pa)
        inverse[:] = list(ia)

    def permutate_bits(self, perm, x):
        pa = (C.c_uint * len(perm))(*perm)
        return self.lib.or_permutate_bits(pa, len(perm), x)

    def inverse_permutate_bits(self, perm, inverse, x):
        pa = (C.c_uint * len(perm))(*perm)
        ia = (C.c_uint * len(inverse))(*inverse)
        return self.lib.or_inverse_permutate_bits(pa, ia, len(perm), x)

    def index_masks(self, qubits):
        q, k = _ua(qubits)
        out = (C.c_uint64 * (k + 1))()
        self.lib.or_make_index_masks(q, k, out)
        return list(out)

    def index_with_qubits(self, g, v, qubits):
        masks = self.index_masks(qubits)
        q, k = _ua(qubits)
        return self.lib.or_index_with_qubits(g, v, q, k, (C.c_uint64 * len(masks))(*masks))

    def inner_product(self, a, b):
        out = (C.c_double * 2)()
        self.lib.or_inner_product(_ptr(a), _ptr(b), a.size, out)
        return complex(out[0], out[1])

    def norm2(self, a):
        return self.lib.or_norm2(_ptr(a), a.size)

    def spin_expectation_value(self, s, qubit):
        out = (C.c_double * 3)()
        self.lib.or_spin_expectation_value(_ptr(s), s.size, qubit, out)
        return tuple(out)

    def expectation_pauli_string(self, s, pauli, qubits):
        q, nq = _ua(qubits)
        out = (C.c_double * 2)()
        self.lib.or_expectation_pauli_string(_ptr(s), s.size, pauli.encode(), q, nq, out)
        return complex(out[0], out[1])

    def expectation_ising(self, s, J, h):
        return self.lib.or_expectation_ising(_ptr(s), s.size, _da(J), _da(h))

    def apply_ising(self, psi, J, h):
        out = np.empty_like(psi)
        self.lib.or_apply_ising(_ptr(out), _ptr(psi), psi.size, _da(J), _da(h))
        return out

    # --- measurement and sampling -------------------------------------------
    def mt64(self, seed):
        """A std::mt19937_64 seeded with `seed` (an object with next() and canonical())."""
        lib = self.lib

        class Mt64:
            def __init__(self):
                self.state = _Mt64State()
                lib.or_mt64_seed(C.byref(self.state), seed)

            def next(self):
                return lib.or_mt64_next(C.byref(self.state))

            def canonical(self):
                return lib.or_mt64_canonical(C.byref(self.state))

        return Mt64()

    def zero_one_probabilities(self, s, qubit):
        out = (C.c_double * 2)()
        self.lib.or_zero_one_probabilities(_ptr(s), s.size, qubit, out)
        return out[0], out[1]

    def project(self, s, qubit, value, multiplier):
        self.lib.or_project(_ptr(s), s.size, qubit, value, multiplier)

    def projective_measurement(self, s, qubit, canonical):
        return self.lib.or_projective_measurement(_ptr(s), s.size, qubit, canonical)

    def generate_events(self, s, canonical):
        canonical = np.ascontiguousarray(canonical, dtype=np.float64)
        out = np.zeros(canonical.size, dtype=np.uint64)
        self.lib.or_generate_events(_ptr(s), s.size, canonical.ctypes.data_as(C.POINTER(C.c_double)), canonical.size,
                                    out.ctypes.data_as(C.POINTER(C.c_uint64)))
        return out

    def shor_box(self, s, m, divisor, base):
        self.lib.or_shor_box(_ptr(s), s.size, m, divisor, base)
