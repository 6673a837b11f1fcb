"""ctypes binding of libbraket_hip.so (the C ABI declared in include/braket_hip.h).

Used by the test and benchmark harness.  Loading fails loudly when the HIP
library has not been built: there is no CPU fallback.

Pointers are plain integers (device addresses, e.g. ``tensor.data_ptr()``);
``stream`` is a ``hipStream_t`` as an integer (``torch.cuda.current_stream().cuda_stream``)
or 0 for the null stream.
"""
from __future__ import annotations

import ctypes as C
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
ROOT = os.path.dirname(PKG)
LIB_PATH = os.path.join(PKG, "libbraket_hip.so")
HEADER = os.path.join(ROOT, "include", "braket_hip.h")

C128, C64 = 0, 1
OK = 0

_u = C.c_uint
_u64 = C.c_uint64
_d = C.c_double
_p = C.c_void_p
_i = C.c_int
_up = C.POINTER(C.c_uint)
_dp = C.POINTER(C.c_double)

# name -> argtypes (all return int status unless listed in _RESTYPES)
_SIGS = {
    "brk_status_string": [_i],
    "brk_last_error": [],
    "brk_version": [],
    "brk_device_count": [C.POINTER(C.c_int)],
    "brk_device_name": [C.c_char_p, C.c_size_t],
    "brk_mem_info": [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)],
    "brk_stream_create": [C.POINTER(C.c_void_p)],
    "brk_stream_destroy": [_p],
    "brk_event_create": [C.POINTER(C.c_void_p)],
    "brk_event_destroy": [_p],
    "brk_event_record": [_p, _p],
    "brk_event_synchronize": [_p],
    "brk_event_elapsed_ms": [_p, _p, C.POINTER(C.c_float)],
    "brk_tune": [C.c_char_p, C.c_int64],
    "brk_get_tune": [C.c_char_p, C.POINTER(C.c_int64)],
    "brk_malloc": [C.POINTER(C.c_void_p), C.c_size_t],
    "brk_free": [_p],
    "brk_memcpy_htod": [_p, _p, C.c_size_t, _p],
    "brk_memcpy_dtoh": [_p, _p, C.c_size_t, _p],
    "brk_memcpy_dtod": [_p, _p, C.c_size_t, _p],
    "brk_stream_synchronize": [_p],
    "brk_set_basis_state": [_p, _u64, _i, _u64, _p],
    "brk_state_create": [_u, _i, _p, C.POINTER(C.c_void_p)],
    "brk_state_destroy": [_p],
    "brk_state_data": [_p, C.POINTER(C.c_void_p), C.POINTER(C.c_uint64), C.POINTER(C.c_int), C.POINTER(C.c_void_p)],
    "brk_state_set_basis": [_p, _u64],
    "brk_state_upload": [_p, _p],
    "brk_state_download": [_p, _p],
    "brk_state_norm2": [_p, _dp],
    "brk_fill_modular": [_p, _u64, _i, _dp, C.POINTER(C.c_uint64), _p],
    "brk_fill_hash": [_p, _u64, _i, _u64, _d, _p],
    "brk_convert": [_p, _i, _p, _i, _u64, _p],
    "brk_copy": [_p, _p, _u64, _i, _p],
    "brk_hadamard": [_p, _u64, _i, _u, _up, _u, _p],
    "brk_pauli_x": [_p, _u64, _i, _up, _u, _up, _u, _p],
    "brk_controlled_not": [_p, _u64, _i, _u, _up, _u, _p],
    "brk_pauli_y": [_p, _u64, _i, _up, _u, _up, _u, _p],
    "brk_pauli_z": [_p, _u64, _i, _up, _u, _p],
    "brk_pauli_zn": [_p, _u64, _i, _up, _u, _up, _u, _p],
    "brk_sqrt_pauli_zn": [_p, _u64, _i, _up, _u, _up, _u, _p],
    "brk_adj_sqrt_pauli_zn": [_p, _u64, _i, _up, _u, _up, _u, _p],
    "brk_phase_shift_coeff": [_p, _u64, _i, _dp, _up, _u, _p],
    "brk_phase_shift": [_p, _u64, _i, _d, _up, _u, _p],
    "brk_phase_shift2": [_p, _u64, _i, _d, _d, _u, _up, _u, _p],
    "brk_phase_shift3": [_p, _u64, _i, _d, _d, _d, _u, _up, _u, _p],
    "brk_exponential_pauli_x": [_p, _u64, _i, _d, _up, _u, _up, _u, _p],
    "brk_exponential_pauli_y": [_p, _u64, _i, _d, _up, _u, _up, _u, _p],
    "brk_exponential_pauli_z": [_p, _u64, _i, _d, _up, _u, _up, _u, _p],
    "brk_swap": [_p, _u64, _i, _u, _u, _up, _u, _p],
    "brk_gate_1q": [_p, _u64, _i, _u, _dp, _up, _u, _p],
    "brk_swap_permutated_local_qubits": [_p, _u64, _i, _u, _u, _p],
    "brk_exponential_swap": [_p, _u64, _i, _d, _u, _u, _up, _u, _p],
    "brk_exponential_swap_coeff": [_p, _u64, _i, _dp, _u, _u, _up, _u, _p],
    "brk_exponential_pauli_x_coeff": [_p, _u64, _i, _dp, _up, _u, _up, _u, _p],
    "brk_exponential_pauli_y_coeff": [_p, _u64, _i, _dp, _up, _u, _up, _u, _p],
    "brk_exponential_pauli_z_coeff": [_p, _u64, _i, _dp, _up, _u, _up, _u, _p],
    "brk_inner_product": [_p, _i, _p, _i, _u64, _dp, _p],
    "brk_fidelity": [_p, _i, _p, _i, _u64, _dp, _p],
    "brk_norm2": [_p, _u64, _i, _dp, _p],
    "brk_spin_expectation_value": [_p, _u64, _i, _u, _dp, _p],
    "brk_expectation_pauli_string": [_p, _u64, _i, C.c_char_p, _up, _u, _dp, _p],
    "brk_expectation_ising": [_p, _u64, _i, _dp, _dp, _dp, _p],
    "brk_apply_ising": [_p, _p, _u64, _i, _dp, _dp, _p],
    "brk_zero_one_probabilities": [_p, _u64, _i, _u, _dp, _p],
    "brk_shor_box": [_p, _u64, _i, _u, _u64, _u64, _p],
    "brk_project_qubit": [_p, _u64, _i, _u, _i, _d, _p],
    "brk_generate_events": [_p, _u64, _i, _dp, _u, C.POINTER(C.c_uint64), _p],
    "brk_generate_events_noalloc": [_p, _u64, _i, _dp, _u, C.POINTER(C.c_uint64), _p],
    "brk_circuit_create": [_u, C.POINTER(C.c_void_p)],
    "brk_circuit_destroy": [_p],
    "brk_circuit_add": [_p, _i, _up, _u, _up, _u, _dp, _u],
    "brk_circuit_num_ops": [_p, _up],
    "brk_circuit_num_qubits": [_p, _up],
    "brk_circuit_run": [_p, _p, _i, _up, _p],
    "brk_circuit_run_timed": [_p, _p, _i, _up, C.POINTER(C.c_float), _p],
    "brk_block_plan_create": [_p, _i, _up, C.POINTER(C.c_void_p)],
    "brk_block_plan_destroy": [_p],
    "brk_block_plan_num_steps": [_p, _up],
    "brk_block_plan_step": [_p, _u, _up, _up, _up, _up],
    "brk_block_plan_final_permutation": [_p, _up],
    "brk_block_plan_run": [_p, _p, _p],
    "brk_circuit_run_blocked": [_p, _p, _i, _up, _p],
    "brk_circuit_block_order": [_p, _u, _u, _up],
    "brk_fused_plan_create": [_p, _u, _up, C.POINTER(C.c_void_p)],
    "brk_fused_plan_destroy": [_p],
    "brk_fused_plan_num_blocks": [_p, _up],
    "brk_fused_plan_num_passes": [_p, _up],
    "brk_fused_plan_block": [_p, _u, _up, _up, _dp],
    "brk_fused_plan_final_permutation": [_p, _up],
    "brk_fused_apply": [_p, _p, _i, _p],
    "brk_fused_apply_timed": [_p, _p, _i, C.POINTER(C.c_float), _p],
    "brk_circuit_num_parameters": [_p, _up],
    "brk_adjoint_ising_gradient": [_p, _p, _p, _i, _dp, _dp, _dp, _dp, _u, _p],
    "brk_circuit_op": [_p, _u, C.POINTER(C.c_int), _up, _up, _up, _up, _dp, _up],
    "brk_qcx_parse": [C.c_char_p, C.POINTER(C.c_void_p)],
    "brk_qcx_destroy": [_p],
    "brk_qcx_num_qubits": [_p, _up],
    "brk_qcx_initial_state": [_p, C.POINTER(_u64)],
    "brk_qcx_num_statements": [_p, _up],
    "brk_qcx_depolarizing": [_p, C.POINTER(C.c_int), _dp, C.POINTER(C.c_int64)],
    "brk_qcx_num_gates": [_p, _u, _up],
    "brk_qcx_gate": [_p, _u, _u, _up, _up, _up, _up],
    "brk_qcx_statement": [_p, _u, C.POINTER(C.c_int), C.POINTER(C.c_void_p), C.POINTER(C.c_int), _up,
                          C.POINTER(C.POINTER(_u64))],
    "brk_batch_create": [C.POINTER(C.c_void_p), _u, C.POINTER(_dp), C.POINTER(_dp), C.POINTER(C.c_void_p)],
    "brk_batch_destroy": [_p],
    "brk_batch_num_stages": [_p, _u, _up],
    "brk_batch_stage": [_p, _u, _u, _up, _up, _up],
    "brk_batch_stage_op": [_p, _u, _u, _u, C.POINTER(C.c_int), C.POINTER(_u64), _dp],
    "brk_batch_run": [_p, _dp, _p],
    "brk_batch_run_timed": [_p, _dp, _dp, _dp, _dp, _p],
}
_RESTYPES = {"brk_status_string": C.c_char_p, "brk_last_error": C.c_char_p, "brk_version": C.c_char_p}

_lib = None


class BrkError(RuntimeError):
    def __init__(self, fn, status, msg):
        super().__init__(f"{fn} failed with status {status}: {msg}")
        self.status = status


def header_symbols():
    """Entry points declared in include/braket_hip.h."""
    with open(HEADER) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(brk_\w+)\s*\(", text)))


def load(path: str = LIB_PATH):
    """Load the engine library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(
            f"{path} not found: build the HIP library first (python -c 'import __graft_entry__ as g; g.build()')")
    lib = C.CDLL(path)
    for name, args in _SIGS.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = _RESTYPES.get(name, C.c_int)
    _lib = lib
    return lib


def _check(name, status):
    if status != OK:
        msg = _lib.brk_last_error().decode()
        raise BrkError(name, status, msg)


def call(name, *args):
    """Call brk_<name> and raise BrkError on a non-zero status."""
    lib = load()
    fn = getattr(lib, name)
    st = fn(*args)
    _check(name, st)
    return st


def uarr(values):
    values = list(values)
    if not values:
        return None, 0
    arr = (C.c_uint * len(values))(*values)
    return arr, len(values)


def darr(values):
    values = list(values)
    return (C.c_double * len(values))(*values)


# ---------------------------------------------------------------------------
# Convenience wrappers (stream defaults to 0 = null stream)
# ---------------------------------------------------------------------------

def hadamard(ptr, size, dtype, target, controls=(), stream=0):
    c, nc = uarr(controls)
    call("brk_hadamard", ptr, size, dtype, target, c, nc, stream)


def pauli_x(ptr, size, dtype, targets, controls=(), stream=0):
    t, nt = uarr(targets)
    c, nc = uarr(controls)
    call("brk_pauli_x", ptr, size, dtype, t, nt, c, nc, stream)


def controlled_not(ptr, size, dtype, target, controls, stream=0):
    c, nc = uarr(controls)
    call("brk_controlled_not", ptr, size, dtype, target, c, nc, stream)


def pauli_y(ptr, size, dtype, targets, controls=(), stream=0):
    t, nt = uarr(targets)
    c, nc = uarr(controls)
    call("brk_pauli_y", ptr, size, dtype, t, nt, c, nc, stream)


def pauli_z(ptr, size, dtype, qubits, stream=0):
    q, nq = uarr(qubits)
    call("brk_pauli_z", ptr, size, dtype, q, nq, stream)


def pauli_zn(ptr, size, dtype, targets, controls=(), stream=0):
    t, nt = uarr(targets)
    c, nc = uarr(controls)
    call("brk_pauli_zn", ptr, size, dtype, t, nt, c, nc, stream)


def sqrt_pauli_zn(ptr, size, dtype, targets, controls=(), stream=0):
    t, nt = uarr(targets)
    c, nc = uarr(controls)
    call("brk_sqrt_pauli_zn", ptr, size, dtype, t, nt, c, nc, stream)


def adj_sqrt_pauli_zn(ptr, size, dtype, targets, controls=(), stream=0):
    t, nt = uarr(targets)
    c, nc = uarr(controls)
    call("brk_adj_sqrt_pauli_zn", ptr, size, dtype, t, nt, c, nc, stream)


def phase_shift_coeff(ptr, size, dtype, coeff, qubits, stream=0):
    q, nq = uarr(qubits)
    call("brk_phase_shift_coeff", ptr, size, dtype, darr([coeff.real, coeff.imag]), q, nq, stream)


def phase_shift(ptr, size, dtype, phase, qubits, stream=0):
    q, nq = uarr(qubits)
    call("brk_phase_shift", ptr, size, dtype, phase, q, nq, stream)


def phase_shift2(ptr, size, dtype, p1, p2, target, controls=(), stream=0):
    c, nc = uarr(controls)
    call("brk_phase_shift2", ptr, size, dtype, p1, p2, target, c, nc, stream)


def phase_shift3(ptr, size, dtype, p1, p2, p3, target, controls=(), stream=0):
    c, nc = uarr(controls)
    call("brk_phase_shift3", ptr, size, dtype, p1, p2, p3, target, c, nc, stream)


def exponential_pauli_x(ptr, size, dtype, phase, targets, controls=(), stream=0):
    t, nt = uarr(targets)
    c, nc = uarr(controls)
    call("brk_exponential_pauli_x", ptr, size, dtype, phase, t, nt, c, nc, stream)


def exponential_pauli_y(ptr, size, dtype, phase, targets, controls=(), stream=0):
    t, nt = uarr(targets)
    c, nc = uarr(controls)
    call("brk_exponential_pauli_y", ptr, size, dtype, phase, t, nt, c, nc, stream)


def exponential_pauli_z(ptr, size, dtype, phase, targets, controls=(), stream=0):
    t, nt = uarr(targets)
    c, nc = uarr(controls)
    call("brk_exponential_pauli_z", ptr, size, dtype, phase, t, nt, c, nc, stream)


def swap(ptr, size, dtype, q1, q2, controls=(), stream=0):
    c, nc = uarr(controls)
    call("brk_swap", ptr, size, dtype, q1, q2, c, nc, stream)


def exponential_swap(ptr, size, dtype, phase, q1, q2, controls=(), stream=0):
    c, nc = uarr(controls)
    call("brk_exponential_swap", ptr, size, dtype, phase, q1, q2, c, nc, stream)


def gate_1q(ptr, size, dtype, target, matrix, controls=(), stream=0):
    m = darr([v for z in matrix for v in (z.real, z.imag)])
    c, nc = uarr(controls)
    call("brk_gate_1q", ptr, size, dtype, target, m, c, nc, stream)


def swap_permutated_local_qubits(ptr, size, dtype, p1, p2, stream=0):
    call("brk_swap_permutated_local_qubits", ptr, size, dtype, p1, p2, stream)


def tune(key, value):
    call("brk_tune", key.encode(), int(value))


def get_tune(key):
    v = C.c_int64()
    call("brk_get_tune", key.encode(), C.byref(v))
    return v.value


def set_basis_state(ptr, size, dtype, index, stream=0):
    call("brk_set_basis_state", ptr, size, dtype, index, stream)


def fill_modular(ptr, size, dtype, scales, params, stream=0):
    call("brk_fill_modular", ptr, size, dtype, darr(scales), (C.c_uint64 * 6)(*params), stream)


def fill_hash(ptr, size, dtype, seed, scale=1.0, stream=0):
    call("brk_fill_hash", ptr, size, dtype, seed, scale, stream)


def inner_product(p1, dt1, p2, dt2, size, stream=0):
    out = (C.c_double * 2)()
    call("brk_inner_product", p1, dt1, p2, dt2, size, out, stream)
    return complex(out[0], out[1])


def norm2(ptr, size, dtype, stream=0):
    out = C.c_double()
    call("brk_norm2", ptr, size, dtype, C.byref(out), stream)
    return out.value


def spin_expectation_value(ptr, size, dtype, qubit, stream=0):
    out = (C.c_double * 3)()
    call("brk_spin_expectation_value", ptr, size, dtype, qubit, out, stream)
    return tuple(out)


def shor_box(ptr, size, dtype, num_exponent_qubits, divisor, base, stream=0):
    call("brk_shor_box", ptr, size, dtype, num_exponent_qubits, divisor, base, stream)


def zero_one_probabilities(ptr, size, dtype, qubit, stream=0):
    out = (C.c_double * 2)()
    call("brk_zero_one_probabilities", ptr, size, dtype, qubit, out, stream)
    return out[0], out[1]


def project_qubit(ptr, size, dtype, qubit, value, multiplier, stream=0):
    call("brk_project_qubit", ptr, size, dtype, qubit, value, multiplier, stream)


def generate_events(ptr, size, dtype, canonical, stream=0, entry="brk_generate_events"):
    """canonical: sequence of numbers in [0, 1); returns a list of basis indices.
    entry="brk_generate_events_noalloc" takes the entry point that allocates
    no device memory."""
    import numpy as np

    c = np.ascontiguousarray(canonical, dtype=np.float64).reshape(-1)
    n = int(c.size)
    out = np.zeros(max(n, 1), dtype=np.uint64)
    call(entry, ptr, size, dtype, c.ctypes.data_as(_dp) if n else None, n,
         out.ctypes.data_as(C.POINTER(C.c_uint64)), stream)
    return out[:n].tolist()


def generate_events_noalloc(ptr, size, dtype, canonical, stream=0):
    return generate_events(ptr, size, dtype, canonical, stream, entry="brk_generate_events_noalloc")


def expectation_pauli_string(ptr, size, dtype, pauli, qubits, stream=0):
    q, nq = uarr(qubits)
    out = (C.c_double * 2)()
    call("brk_expectation_pauli_string", ptr, size, dtype, pauli.encode(), q, nq, out, stream)
    return complex(out[0], out[1])


def expectation_ising(ptr, size, dtype, J, h, stream=0):
    out = C.c_double()
    call("brk_expectation_ising", ptr, size, dtype, darr(J), darr(h), C.byref(out), stream)
    return out.value


def apply_ising(out_ptr, psi_ptr, size, dtype, J, h, stream=0):
    call("brk_apply_ising", out_ptr, psi_ptr, size, dtype, darr(J), darr(h), stream)


def copy(dst, src, size, dtype, stream=0):
    call("brk_copy", dst, src, size, dtype, stream)


def convert(dst, dt_dst, src, dt_src, size, stream=0):
    call("brk_convert", dst, dt_dst, src, dt_src, size, stream)


# ---------------------------------------------------------------------------
# Circuits
# ---------------------------------------------------------------------------

OP = dict(H=0, X=1, Y=2, Z=3, U1=4, U2=5, U3=6, EX=7, EY=8, EZ=9, RX=10, RY=11, RZ=12, SWAP=13,
          LOCAL_SWAP=14, MATRIX=15, ESWAP=16, PHASE=17)


class BlockPlan:
    """brk_block_plan of a Circuit: steps are (first_op, num_ops, local_qubits), local_qubits empty for a
    plain step."""

    def __init__(self, circuit: "Circuit", dtype=C128, permutation=None):
        self.circuit = circuit
        n = circuit.num_qubits
        pa = None if permutation is None else (C.c_uint * n)(*permutation)
        h = C.c_void_p()
        call("brk_block_plan_create", circuit.h, dtype, pa, C.byref(h))
        self.h = h.value
        cnt = C.c_uint()
        call("brk_block_plan_num_steps", self.h, C.byref(cnt))
        self.steps = []
        for i in range(cnt.value):
            f, k, nl = C.c_uint(), C.c_uint(), C.c_uint()
            lq = (C.c_uint * 16)()
            call("brk_block_plan_step", self.h, i, C.byref(f), C.byref(k), C.byref(nl), lq)
            self.steps.append((f.value, k.value, tuple(lq[: nl.value])))
        fp = (C.c_uint * n)()
        call("brk_block_plan_final_permutation", self.h, fp)
        self.final_permutation = list(fp)

    def run(self, ptr, stream=0):
        call("brk_block_plan_run", self.h, ptr, stream)

    def __del__(self):
        try:
            if self.h:
                load().brk_block_plan_destroy(self.h)
                self.h = None
        except Exception:
            pass


class Circuit:
    """Owns a brk_circuit built from (name, targets, controls, params) tuples."""

    def __init__(self, num_qubits, ops=()):
        h = C.c_void_p()
        call("brk_circuit_create", num_qubits, C.byref(h))
        self.h = h.value
        self.num_qubits = num_qubits
        self.ops = []
        for op in ops:
            self.add(*op)

    def add(self, name, targets, controls=(), params=()):
        t, nt = uarr(targets)
        c, nc = uarr(controls)
        p = darr(params) if len(params) else None
        call("brk_circuit_add", self.h, OP[name], t, nt, c, nc, p, len(params))
        self.ops.append((name, tuple(targets), tuple(controls), tuple(params)))

    def __len__(self):
        return len(self.ops)

    def run(self, ptr, dtype=C128, permutation=None, stream=0):
        """Runs on the device state; returns the final logical->physical permutation."""
        perm = list(range(self.num_qubits)) if permutation is None else list(permutation)
        pa = (C.c_uint * self.num_qubits)(*perm)
        call("brk_circuit_run", self.h, ptr, dtype, pa, stream)
        return list(pa)

    def run_blocked(self, ptr, dtype=C128, permutation=None, stream=0):
        """brk_circuit_run_blocked: gate blocks in LDS, same amplitudes as run()."""
        perm = list(range(self.num_qubits)) if permutation is None else list(permutation)
        pa = (C.c_uint * self.num_qubits)(*perm)
        call("brk_circuit_run_blocked", self.h, ptr, dtype, pa, stream)
        return list(pa)

    def block_order(self, num_local, line_bits):
        """brk_circuit_block_order: op indices in an order that packs into fewer gate blocks."""
        out = (C.c_uint * max(1, len(self.ops)))()
        call("brk_circuit_block_order", self.h, num_local, line_bits, out)
        return list(out)[: len(self.ops)]

    def reordered(self, order):
        """The same circuit with its ops in `order` (a list of op indices)."""
        return Circuit(self.num_qubits, [self.ops[i] for i in order])

    def run_timed(self, ptr, dtype=C128, permutation=None, stream=0):
        perm = list(range(self.num_qubits)) if permutation is None else list(permutation)
        pa = (C.c_uint * self.num_qubits)(*perm)
        ms = (C.c_float * max(1, len(self.ops)))()
        call("brk_circuit_run_timed", self.h, ptr, dtype, pa, ms, stream)
        return list(pa), list(ms)[: len(self.ops)]

    def __del__(self):
        try:
            if self.h:
                load().brk_circuit_destroy(self.h)
                self.h = None
        except Exception:
            pass


class FusedPlan:
    """brk_fused_plan of a Circuit (precision form)."""

    def __init__(self, circuit: "Circuit", width: int, permutation=None):
        self.num_qubits = circuit.num_qubits
        self.width = width
        pa = None if permutation is None else (C.c_uint * self.num_qubits)(*permutation)
        h = C.c_void_p()
        call("brk_fused_plan_create", circuit.h, width, pa, C.byref(h))
        self.h = h.value
        n = C.c_uint()
        call("brk_fused_plan_num_blocks", self.h, C.byref(n))
        self.num_blocks = n.value
        call("brk_fused_plan_num_passes", self.h, C.byref(n))
        self.num_passes = n.value

    def block(self, i, with_matrix=True):
        import numpy as np
        q = (C.c_uint * self.width)()
        g = C.c_uint()
        d = 1 << self.width
        m = (C.c_double * (2 * d * d))() if with_matrix else None
        call("brk_fused_plan_block", self.h, i, q, C.byref(g), m)
        mat = None
        if with_matrix:
            a = np.frombuffer(m, dtype=np.float64)
            mat = (a[0::2] + 1j * a[1::2]).reshape(d, d)
        return list(q), g.value, mat

    def final_permutation(self):
        p = (C.c_uint * self.num_qubits)()
        call("brk_fused_plan_final_permutation", self.h, p)
        return list(p)

    def apply(self, ptr, dtype=C64, stream=0):
        call("brk_fused_apply", self.h, ptr, dtype, stream)

    def apply_timed(self, ptr, dtype=C64, stream=0):
        ms = (C.c_float * max(1, self.num_blocks))()
        call("brk_fused_apply_timed", self.h, ptr, dtype, ms, stream)
        return list(ms)[: self.num_blocks]

    def __del__(self):
        try:
            if self.h:
                load().brk_fused_plan_destroy(self.h)
                self.h = None
        except Exception:
            pass


def circuit_num_parameters(circuit):
    n = C.c_uint()
    call("brk_circuit_num_parameters", circuit.h, C.byref(n))
    return n.value


def adjoint_ising_gradient(circuit, psi_ptr, work_ptr, dtype, J, h, stream=0):
    """(energy, gradient list) of <psi|H|psi>; psi_ptr holds |psi0> on entry."""
    import numpy as np
    npar = circuit_num_parameters(circuit)
    e = C.c_double()
    g = (C.c_double * max(1, npar))()
    call("brk_adjoint_ising_gradient", circuit.h, psi_ptr, work_ptr, dtype, darr(J), darr(h), C.byref(e), g, npar,
         stream)
    return e.value, np.array(list(g)[:npar])


BATCH_GATE, BATCH_SIGN, BATCH_MEASURE_X, BATCH_MEASURE_ZZ = 0, 1, 2, 3


class Batch:
    """brk_batch over [(num_qubits, ops, J, h)] (batching form)."""

    def __init__(self, items):
        self.circuits = [Circuit(n, ops) for n, ops, _, _ in items]
        self.num_circuits = len(items)
        self._J = [darr(J) if J is not None else None for _, _, J, _ in items]
        self._h = [darr(h) if h is not None else None for _, _, _, h in items]
        handles = (C.c_void_p * self.num_circuits)(*[c.h for c in self.circuits])
        Ja = (_dp * self.num_circuits)(*[C.cast(a, _dp) if a is not None else None for a in self._J])
        ha = (_dp * self.num_circuits)(*[C.cast(a, _dp) if a is not None else None for a in self._h])
        out = C.c_void_p()
        call("brk_batch_create", handles, self.num_circuits, Ja, ha, C.byref(out))
        self.h = out.value

    def num_stages(self, circuit):
        n = C.c_uint()
        call("brk_batch_num_stages", self.h, circuit, C.byref(n))
        return n.value

    def stage(self, circuit, stage):
        """(local qubits, [(kind, qubit mask, 2x2 matrix or None)])."""
        import numpy as np
        local = (C.c_uint * 32)()
        nl, nops = C.c_uint(), C.c_uint()
        call("brk_batch_stage", self.h, circuit, stage, local, C.byref(nl), C.byref(nops))
        ops = []
        for k in range(nops.value):
            kind, mask = C.c_int(), C.c_uint64()
            m = (C.c_double * 8)()
            call("brk_batch_stage_op", self.h, circuit, stage, k, C.byref(kind), C.byref(mask), m)
            a = np.frombuffer(m, dtype=np.float64)
            mat = (a[0::2] + 1j * a[1::2]).reshape(2, 2) if kind.value == BATCH_GATE else None
            ops.append((kind.value, mask.value, mat))
        return list(local)[: nl.value], ops

    def run(self, stream=0):
        import numpy as np
        e = (C.c_double * self.num_circuits)()
        call("brk_batch_run", self.h, e, stream)
        return np.array(list(e))

    def run_timed(self, stream=0):
        """(energies, graph ms, per-circuit start ms, per-circuit end ms)."""
        import numpy as np
        nc = self.num_circuits
        e, t0, t1 = (C.c_double * nc)(), (C.c_double * nc)(), (C.c_double * nc)()
        ms = C.c_double()
        call("brk_batch_run_timed", self.h, e, C.byref(ms), t0, t1, stream)
        return np.array(list(e)), ms.value, np.array(list(t0)), np.array(list(t1))

    def __del__(self):
        try:
            if self.h:
                load().brk_batch_destroy(self.h)
                self.h = None
        except Exception:
            pass


OP_NAMES = ("H", "X", "Y", "Z", "U1", "U2", "U3", "EX", "EY", "EZ", "RX", "RY", "RZ", "SWAP", "LOCAL_SWAP", "MATRIX",
            "ESWAP", "PHASE")


def circuit_ops(handle):
    """The ops of a brk_circuit handle as (name, targets, controls, params) tuples."""
    n = C.c_uint()
    call("brk_circuit_num_ops", handle, C.byref(n))
    out = []
    for i in range(n.value):
        code, nt, nc, npar = C.c_int(), C.c_uint(), C.c_uint(), C.c_uint()
        t, c, p = (C.c_uint * 16)(), (C.c_uint * 16)(), (C.c_double * 8)()
        call("brk_circuit_op", handle, i, C.byref(code), t, C.byref(nt), c, C.byref(nc), p, C.byref(npar))
        out.append((OP_NAMES[code.value], tuple(t[: nt.value]), tuple(c[: nc.value]), tuple(p[: npar.value])))
    return out


QCX_GATES, QCX_MEASUREMENT, QCX_AMPLITUDES, QCX_BEGIN_FUSION, QCX_END_FUSION = 0, 1, 2, 3, 4
QCX_PROJECTIVE_MEASUREMENT, QCX_CLEAR, QCX_SET, QCX_GENERATE_EVENTS, QCX_EXIT = 5, 6, 7, 8, 9
QCX_SHOR_BOX = 10


class Qcx:
    """A parsed quantum-assembler program (bra nompi subset)."""

    def __init__(self, text):
        h = C.c_void_p()
        call("brk_qcx_parse", text.encode(), C.byref(h))
        self.h = h.value
        n, init, cnt = C.c_uint(), C.c_uint64(), C.c_uint()
        call("brk_qcx_num_qubits", self.h, C.byref(n))
        call("brk_qcx_initial_state", self.h, C.byref(init))
        call("brk_qcx_num_statements", self.h, C.byref(cnt))
        self.num_qubits, self.initial_state = n.value, init.value
        self.statements = []
        self.gate_records = {}  # statement -> [(first_op, num_ops, noise qubits)] per program gate
        en, probs, seed = C.c_int(), (C.c_double * 3)(), C.c_int64()
        call("brk_qcx_depolarizing", self.h, C.byref(en), probs, C.byref(seed))
        self.depolarizing = (list(probs), seed.value) if en.value else None
        for i in range(cnt.value):
            kind, prec, nidx = C.c_int(), C.c_int(), C.c_uint()
            gates = C.c_void_p()
            idx = C.POINTER(_u64)()
            call("brk_qcx_statement", self.h, i, C.byref(kind), C.byref(gates), C.byref(prec), C.byref(nidx),
                 C.byref(idx))
            if kind.value == QCX_GATES:
                self.statements.append((kind.value, circuit_ops(gates.value)))
                ng = C.c_uint()
                call("brk_qcx_num_gates", self.h, i, C.byref(ng))
                records = []
                for g in range(ng.value):
                    first, count, nn = C.c_uint(), C.c_uint(), C.c_uint()
                    noise = (C.c_uint * 16)()
                    call("brk_qcx_gate", self.h, i, g, C.byref(first), C.byref(count), noise, C.byref(nn))
                    records.append((first.value, count.value, list(noise[: nn.value])))
                self.gate_records[i] = records
            elif kind.value == QCX_MEASUREMENT:
                self.statements.append((kind.value, prec.value))
            elif kind.value == QCX_AMPLITUDES:
                self.statements.append((kind.value, [idx[k] for k in range(nidx.value)]))
            elif kind.value in (QCX_PROJECTIVE_MEASUREMENT, QCX_CLEAR, QCX_SET):
                self.statements.append((kind.value, idx[0]))
            elif kind.value == QCX_SHOR_BOX:
                self.statements.append((kind.value, (idx[0], idx[1], idx[2])))  # (m, divisor, base)
            elif kind.value == QCX_GENERATE_EVENTS:
                self.statements.append((kind.value, (idx[0], prec.value)))  # (num_events, seed)
            else:
                self.statements.append((kind.value, None))

    def ops(self):
        """Every gate of the program in order."""
        return [op for kind, v in self.statements if kind == QCX_GATES for op in v]

    def __del__(self):
        try:
            if self.h:
                load().brk_qcx_destroy(self.h)
                self.h = None
        except Exception:
            pass


def device_name():
    buf = C.create_string_buffer(256)
    call("brk_device_name", buf, 256)
    return buf.value.decode()
