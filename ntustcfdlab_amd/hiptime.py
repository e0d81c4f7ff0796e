"""HIP events on a given stream, through the HIP runtime's C API (ctypes).

Used by bench.py to time work enqueued by libntust_hip.so on its own stream.
"""
from __future__ import annotations

import ctypes as ct

_hip = None


def _rt():
    global _hip
    if _hip is None:
        for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
            try:
                _hip = ct.CDLL(name)
                break
            except OSError:
                continue
        if _hip is None:
            raise OSError("libamdhip64.so not found")
        _hip.hipEventCreate.argtypes = [ct.POINTER(ct.c_void_p)]
        _hip.hipEventRecord.argtypes = [ct.c_void_p, ct.c_void_p]
        _hip.hipEventSynchronize.argtypes = [ct.c_void_p]
        _hip.hipEventElapsedTime.argtypes = [ct.POINTER(ct.c_float), ct.c_void_p, ct.c_void_p]
        _hip.hipEventDestroy.argtypes = [ct.c_void_p]
        _hip.hipDeviceSynchronize.argtypes = []
    return _hip


def _ok(rc):
    if rc != 0:
        raise RuntimeError(f"HIP runtime call failed with {rc}")


class Event:
    def __init__(self):
        self.h = ct.c_void_p()
        _ok(_rt().hipEventCreate(ct.byref(self.h)))

    def record(self, stream):
        _ok(_rt().hipEventRecord(self.h, ct.c_void_p(stream)))

    def elapsed_ms(self, later: "Event") -> float:
        _ok(_rt().hipEventSynchronize(later.h))
        ms = ct.c_float()
        _ok(_rt().hipEventElapsedTime(ct.byref(ms), self.h, later.h))
        return ms.value

    def __del__(self):
        try:
            _rt().hipEventDestroy(self.h)
        except Exception:
            pass


def device_synchronize():
    _ok(_rt().hipDeviceSynchronize())


def copy_calibration(nbytes: int = 256 << 20) -> int:
    """One device-to-device copy of `nbytes` (the runtime's copy kernel), so a
    counter profile of the same process holds a plain copy of known size to
    scale FETCH_SIZE / WRITE_SIZE by.  Returns nbytes."""
    rt = _rt()
    rt.hipMalloc.argtypes = [ct.POINTER(ct.c_void_p), ct.c_size_t]
    rt.hipFree.argtypes = [ct.c_void_p]
    rt.hipMemset.argtypes = [ct.c_void_p, ct.c_int, ct.c_size_t]
    rt.hipMemcpy.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_size_t, ct.c_int]
    a, b = ct.c_void_p(), ct.c_void_p()
    _ok(rt.hipMalloc(ct.byref(a), nbytes))
    _ok(rt.hipMalloc(ct.byref(b), nbytes))
    _ok(rt.hipMemset(a, 1, nbytes))
    _ok(rt.hipDeviceSynchronize())
    _ok(rt.hipMemcpy(b, a, nbytes, 3))   # hipMemcpyDeviceToDevice
    _ok(rt.hipDeviceSynchronize())
    _ok(rt.hipFree(a))
    _ok(rt.hipFree(b))
    return nbytes
