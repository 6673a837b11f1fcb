"""Minimal ctypes access to the HIP runtime for the harness: events (kernel
timing on the stream the engine launches on), streams and device queries."""
from __future__ import annotations

import ctypes as C

_hip = None


def lib():
    global _hip
    if _hip is None:
        _hip = C.CDLL("libamdhip64.so")
        _hip.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
        _hip.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        _hip.hipEventSynchronize.argtypes = [C.c_void_p]
        _hip.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
        _hip.hipEventDestroy.argtypes = [C.c_void_p]
        _hip.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
        _hip.hipStreamDestroy.argtypes = [C.c_void_p]
        _hip.hipStreamSynchronize.argtypes = [C.c_void_p]
        _hip.hipDeviceGetName.argtypes = [C.c_char_p, C.c_int, C.c_int]
        _hip.hipMemGetInfo.argtypes = [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]
    return _hip


def _check(name, st):
    if st != 0:
        raise RuntimeError(f"{name} failed with hipError {st}")


class Event:
    def __init__(self):
        self.h = C.c_void_p()
        _check("hipEventCreate", lib().hipEventCreate(C.byref(self.h)))

    def record(self, stream=0):
        _check("hipEventRecord", lib().hipEventRecord(self.h, stream))

    def synchronize(self):
        _check("hipEventSynchronize", lib().hipEventSynchronize(self.h))

    def elapsed_ms(self, end: "Event") -> float:
        ms = C.c_float()
        _check("hipEventElapsedTime", lib().hipEventElapsedTime(C.byref(ms), self.h, end.h))
        return ms.value

    def __del__(self):
        try:
            if self.h:
                lib().hipEventDestroy(self.h)
        except Exception:
            pass


def create_stream():
    s = C.c_void_p()
    _check("hipStreamCreate", lib().hipStreamCreate(C.byref(s)))
    return s.value


def synchronize(stream=0):
    _check("hipStreamSynchronize", lib().hipStreamSynchronize(stream))


def mem_info():
    free, total = C.c_size_t(), C.c_size_t()
    _check("hipMemGetInfo", lib().hipMemGetInfo(C.byref(free), C.byref(total)))
    return free.value, total.value
