"""Minimal ctypes access to the HIP runtime for host-side plumbing: device
buffers (to keep D and V resident in HBM) and events on a given stream."""
import ctypes

import numpy as np

_hip = None


def hip():
    global _hip
    if _hip is None:
        L = ctypes.CDLL("libamdhip64.so")
        vp, i, sz = ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t
        L.hipMalloc.argtypes = [ctypes.POINTER(vp), sz]
        L.hipFree.argtypes = [vp]
        L.hipMemcpy.argtypes = [vp, vp, sz, i]
        L.hipDeviceSynchronize.argtypes = []
        L.hipStreamSynchronize.argtypes = [vp]
        L.hipEventCreate.argtypes = [ctypes.POINTER(vp)]
        L.hipEventRecord.argtypes = [vp, vp]
        L.hipEventSynchronize.argtypes = [vp]
        L.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), vp, vp]
        L.hipEventDestroy.argtypes = [vp]
        L.hipGetErrorString.restype = ctypes.c_char_p
        L.hipGetErrorString.argtypes = [i]
        _hip = L
    return _hip


def _ok(rc, what):
    if rc != 0:
        raise RuntimeError("%s: %s" % (what, hip().hipGetErrorString(rc).decode()))


class DeviceArray:
    """A float64 buffer in device memory."""

    def __init__(self, n):
        self.n = int(n)
        p = ctypes.c_void_p()
        _ok(hip().hipMalloc(ctypes.byref(p), max(8, 8 * self.n)), "hipMalloc")
        self.ptr = p.value

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64).ravel(order="K")
        assert a.size == self.n
        _ok(hip().hipMemcpy(self.ptr, a.ctypes.data, 8 * self.n, 1), "hipMemcpy H2D")

    def download(self):
        out = np.empty(self.n)
        _ok(hip().hipMemcpy(out.ctypes.data, self.ptr, 8 * self.n, 2), "hipMemcpy D2H")
        return out

    def __del__(self):
        if getattr(self, "ptr", None) and _hip is not None:
            _hip.hipFree(self.ptr)
            self.ptr = None


class Event:
    def __init__(self):
        e = ctypes.c_void_p()
        _ok(hip().hipEventCreate(ctypes.byref(e)), "hipEventCreate")
        self.e = e.value

    def record(self, stream):
        _ok(hip().hipEventRecord(self.e, stream), "hipEventRecord")

    def elapsed_ms(self, later):
        hip().hipEventSynchronize(later.e)
        ms = ctypes.c_float()
        _ok(hip().hipEventElapsedTime(ctypes.byref(ms), self.e, later.e), "hipEventElapsedTime")
        return ms.value

    def __del__(self):
        if getattr(self, "e", None) and _hip is not None:
            _hip.hipEventDestroy(self.e)


def copy_device(dst, src, nbytes):
    """Device-to-device copy of nbytes (the runtime's copy kernel): a known-size streaming probe."""
    _ok(hip().hipMemcpy(dst, src, int(nbytes), 3), "hipMemcpy D2D")


def device_synchronize():
    _ok(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")
