"""GPU checks of the device-state handle (brk_state_*): a state created at |0..0>, gates run on
its (first, size, dtype, stream) through the plain entry points, whole-state copies both ways,
|basis> preparation and the norm, for both dtypes; H on qubit q against the closed form."""
import ctypes as C
import sys

import numpy as np
import pytest

import brk
import dense as D

pytestmark = pytest.mark.gpu


def note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("dtype", [brk.C128, brk.C64])
def test_state_handle_roundtrip(dtype):
    n = 10
    note(f"gpu state handle dtype={dtype}")
    np_t = np.complex128 if dtype == brk.C128 else np.complex64
    h = C.c_void_p()
    brk.call("brk_state_create", n, dtype, None, C.byref(h))
    try:
        first, size, dt, stream = C.c_void_p(), C.c_uint64(), C.c_int(), C.c_void_p()
        brk.call("brk_state_data", h, C.byref(first), C.byref(size), C.byref(dt), C.byref(stream))
        assert size.value == 1 << n and dt.value == dtype and stream.value is None
        out = np.empty(1 << n, dtype=np_t)
        brk.call("brk_state_download", h, out.ctypes.data)
        expected = np.zeros(1 << n, dtype=np_t)
        expected[0] = 1
        assert np.array_equal(out, expected)
        # H on qubit 3 of |0..0>: amplitudes 1/sqrt2 at 0 and 8
        brk.hadamard(first.value, size.value, dtype, 3)
        brk.call("brk_state_download", h, out.ctypes.data)
        ref = np.zeros(1 << n, dtype=np.complex128)
        ref[0] = ref[8] = 2 ** -0.5
        assert np.max(np.abs(out - ref)) < (1e-15 if dtype == brk.C128 else 1e-7)
        # upload a random state, read it back bit for bit, check its norm
        s = D.random_state(n, np.random.default_rng(4)).astype(np_t)
        brk.call("brk_state_upload", h, s.ctypes.data)
        brk.call("brk_state_download", h, out.ctypes.data)
        assert np.array_equal(out, s)
        nrm = C.c_double()
        brk.call("brk_state_norm2", h, C.byref(nrm))
        assert abs(nrm.value - float(np.vdot(s, s).real)) < (1e-12 if dtype == brk.C128 else 1e-6)
        # |5>
        brk.call("brk_state_set_basis", h, 5)
        brk.call("brk_state_download", h, out.ctypes.data)
        expected = np.zeros(1 << n, dtype=np_t)
        expected[5] = 1
        assert np.array_equal(out, expected)
    finally:
        brk.call("brk_state_destroy", h)
