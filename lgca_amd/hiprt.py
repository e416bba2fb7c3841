"""Minimal ctypes access to the HIP runtime that liblgca_hip.so is linked
against: streams, events and copies for benchmarks and tests."""
from __future__ import annotations

import ctypes as C

from ._capi import lib

_hip = None

MEMCPY_H2D, MEMCPY_D2H, MEMCPY_D2D = 1, 2, 3


def hip() -> C.CDLL:
    global _hip
    if _hip is None:
        lib()  # loads the engine, which pulls in its HIP runtime
        h = C.CDLL("libamdhip64.so.7")
        h.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
        h.hipStreamDestroy.argtypes = [C.c_void_p]
        h.hipStreamSynchronize.argtypes = [C.c_void_p]
        h.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
        h.hipEventDestroy.argtypes = [C.c_void_p]
        h.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        h.hipEventSynchronize.argtypes = [C.c_void_p]
        h.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
        h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        h.hipFree.argtypes = [C.c_void_p]
        h.hipDeviceSynchronize.argtypes = []
        _hip = h
    return _hip


def _ok(code: int, what: str) -> None:
    if code != 0:
        raise RuntimeError(f"{what} failed with HIP error {code}")


class Stream:
    def __init__(self):
        self.ptr = C.c_void_p()
        _ok(hip().hipStreamCreate(C.byref(self.ptr)), "hipStreamCreate")

    def synchronize(self) -> None:
        _ok(hip().hipStreamSynchronize(self.ptr), "hipStreamSynchronize")

    def close(self) -> None:
        if self.ptr:
            hip().hipStreamDestroy(self.ptr)
            self.ptr = C.c_void_p()


class Event:
    def __init__(self):
        self.ptr = C.c_void_p()
        _ok(hip().hipEventCreate(C.byref(self.ptr)), "hipEventCreate")

    def record(self, stream: Stream) -> None:
        _ok(hip().hipEventRecord(self.ptr, stream.ptr), "hipEventRecord")

    def synchronize(self) -> None:
        _ok(hip().hipEventSynchronize(self.ptr), "hipEventSynchronize")

    def elapsed_ms(self, end: "Event") -> float:
        ms = C.c_float()
        _ok(hip().hipEventElapsedTime(C.byref(ms), self.ptr, end.ptr), "hipEventElapsedTime")
        return float(ms.value)

    def close(self) -> None:
        if self.ptr:
            hip().hipEventDestroy(self.ptr)
            self.ptr = C.c_void_p()


def memcpy_d2h(dst_host_ptr: int, src_dev_ptr: int, nbytes: int) -> None:
    _ok(hip().hipMemcpy(dst_host_ptr, src_dev_ptr, nbytes, MEMCPY_D2H), "hipMemcpy")
