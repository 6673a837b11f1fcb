"""Streams, events and device-memory queries for the harness, through
libbraket_hip.so itself: the process then holds a single HIP runtime (the
one the engine was linked against), whatever else (e.g. torch) is loaded."""
from __future__ import annotations

import ctypes as C

import brk


class Event:
    def __init__(self):
        self.h = C.c_void_p()
        brk.call("brk_event_create", C.byref(self.h))

    def record(self, stream=0):
        brk.call("brk_event_record", self.h, stream)

    def synchronize(self):
        brk.call("brk_event_synchronize", self.h)

    def elapsed_ms(self, end: "Event") -> float:
        ms = C.c_float()
        brk.call("brk_event_elapsed_ms", self.h, end.h, C.byref(ms))
        return ms.value

    def __del__(self):
        try:
            if self.h:
                brk.load().brk_event_destroy(self.h)
                self.h = None
        except Exception:
            pass


def create_stream():
    s = C.c_void_p()
    brk.call("brk_stream_create", C.byref(s))
    return s.value


def synchronize(stream=0):
    brk.call("brk_stream_synchronize", stream)


def mem_info():
    free, total = C.c_size_t(), C.c_size_t()
    brk.call("brk_mem_info", C.byref(free), C.byref(total))
    return free.value, total.value
