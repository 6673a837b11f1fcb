"""CPU checks of the C ABI boundary: the HIP library loads and exports every
entry point include/braket_hip.h declares, the binding covers them, and
argument errors are reported as status codes without touching a GPU."""
import ctypes as C
import subprocess

import pytest

import brk


def test_library_exports_every_declared_symbol():
    declared = brk.header_symbols()
    assert len(declared) > 30
    out = subprocess.run(["nm", "-D", "--defined-only", brk.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    missing = [s for s in declared if s not in exported]
    assert not missing, missing


def test_binding_covers_header():
    lib = brk.load()
    for name in brk.header_symbols():
        assert hasattr(lib, name)
        assert name in brk._SIGS, f"{name} has no ctypes signature"


def test_status_strings():
    lib = brk.load()
    assert lib.brk_status_string(0) == b"ok"
    assert lib.brk_status_string(2) == b"qubit out of range"


# Validation happens on the host before any HIP call, so these run without a GPU.
# The pointer is a dummy non-null value that is never dereferenced.
DUMMY = 0x1000


def _status(fn, *args):
    lib = brk.load()
    return getattr(lib, fn)(*args)


def test_argument_errors_without_gpu():
    u = brk.uarr
    t, nt = u([1])
    c, nc = u([1])
    assert _status("brk_hadamard", None, 16, 0, 0, None, 0, None) == 1  # null state
    assert _status("brk_hadamard", DUMMY, 12, 0, 0, None, 0, None) == 1  # not a power of two
    assert _status("brk_hadamard", DUMMY, 16, 7, 0, None, 0, None) == 5  # bad dtype
    assert _status("brk_hadamard", DUMMY, 16, 0, 4, None, 0, None) == 2  # qubit 4 of 4 qubits
    assert _status("brk_hadamard", DUMMY, 16, 0, 1, c, nc, None) == 3  # target == control
    assert _status("brk_swap", DUMMY, 16, 0, 2, 2, None, 0, None) == 3
    many, nm = u(range(20))
    assert _status("brk_pauli_x", DUMMY, 1 << 30, 0, many, nm, None, 0, None) == 4
    assert _status("brk_pauli_x", DUMMY, 16, 0, None, 0, None, 0, None) == 1  # no targets
    assert brk.load().brk_last_error() != b""


def test_block_order_argument_errors():
    circuit = brk.Circuit(4, [("H", (0,), (), ()), ("X", (1,), (0,), ())])
    out = (brk.C.c_uint * 2)()
    assert _status("brk_circuit_block_order", None, 12, 3, out) == 1  # null circuit
    assert _status("brk_circuit_block_order", circuit.h, 12, 3, None) == 1  # null order, ops present
    assert _status("brk_circuit_block_order", circuit.h, 0, 0, out) == 1  # no local qubits
    assert _status("brk_circuit_block_order", circuit.h, 65, 3, out) == 1  # beyond 64 bits
    assert _status("brk_circuit_block_order", circuit.h, 3, 4, out) == 1  # line bits above num_local
    assert _status("brk_circuit_block_order", circuit.h, 12, 3, out) == 0
    assert list(out) == [0, 1]
    empty = brk.Circuit(3)
    assert _status("brk_circuit_block_order", empty.h, 12, 3, None) == 0  # nothing to order


def test_state_handle_argument_errors():
    h = C.c_void_p()
    assert _status("brk_state_create", 10, 7, None, C.byref(h)) == 5  # unknown dtype
    assert _status("brk_state_create", 41, 0, None, C.byref(h)) == 1  # beyond one device
    assert _status("brk_state_create", 10, 0, None, None) == 1
    assert _status("brk_state_data", None, None, None, None, None) == 1
    assert _status("brk_state_set_basis", None, 0) == 1
    assert _status("brk_state_norm2", None, None) == 1
    assert _status("brk_state_destroy", None) == 0
