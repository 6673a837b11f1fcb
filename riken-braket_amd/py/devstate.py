"""Device state vectors for the harness: memory from brk_malloc, copies through the C ABI."""
from __future__ import annotations

import ctypes as C

import numpy as np

import brk

_NP = {brk.C128: np.complex128, brk.C64: np.complex64}
_BYTES = {brk.C128: 16, brk.C64: 8}


class DeviceState:
    """2^n amplitudes in HBM.  Freed explicitly (free()) or when garbage collected."""

    def __init__(self, num_qubits: int, dtype: int = brk.C128, host: np.ndarray | None = None):
        self.num_qubits = num_qubits
        self.size = 1 << num_qubits
        self.dtype = dtype
        self.nbytes = self.size * _BYTES[dtype]
        ptr = C.c_void_p()
        brk.call("brk_malloc", C.byref(ptr), self.nbytes)
        self.ptr = ptr.value
        if host is not None:
            self.upload(host)

    @classmethod
    def from_host(cls, arr: np.ndarray, dtype: int = brk.C128):
        n = int(arr.size).bit_length() - 1
        assert arr.size == 1 << n
        return cls(n, dtype, arr)

    def upload(self, arr: np.ndarray):
        a = np.ascontiguousarray(arr, dtype=_NP[self.dtype])
        assert a.size == self.size
        brk.call("brk_memcpy_htod", self.ptr, a.ctypes.data, self.nbytes, None)

    def download(self) -> np.ndarray:
        out = np.empty(self.size, dtype=_NP[self.dtype])
        brk.call("brk_memcpy_dtoh", out.ctypes.data, self.ptr, self.nbytes, None)
        return out

    def download_indices(self, indices) -> np.ndarray:
        """Gather a few amplitudes (one small copy per index) without moving the whole state."""
        out = np.empty(len(indices), dtype=_NP[self.dtype])
        b = _BYTES[self.dtype]
        for k, i in enumerate(indices):
            brk.call("brk_memcpy_dtoh", out[k:].ctypes.data, self.ptr + int(i) * b, b, None)
        return out

    def free(self):
        if self.ptr:
            brk.call("brk_free", self.ptr)
            self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
