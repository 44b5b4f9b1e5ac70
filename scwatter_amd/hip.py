"""Minimal HIP runtime plumbing over ctypes: device memory, copies, streams,
events.  Uses the libamdhip64 that libscwatter_amd.so itself links, so device
pointers and streams are shared with the library.  Nothing here computes.
"""
import ctypes

import numpy as np

from . import _native

_hip = None
_vp = ctypes.c_void_p

H2D, D2H, D2D = 1, 2, 3


def hip():
    global _hip
    if _hip is not None:
        return _hip
    _native.lib()  # loads libamdhip64.so.7 as a dependency first
    h = ctypes.CDLL("libamdhip64.so.7")
    sig = {
        "hipMalloc": [ctypes.POINTER(_vp), ctypes.c_size_t],
        "hipFree": [_vp],
        "hipMemcpy": [_vp, _vp, ctypes.c_size_t, ctypes.c_int],
        "hipMemcpyAsync": [_vp, _vp, ctypes.c_size_t, ctypes.c_int, _vp],
        "hipMemsetAsync": [_vp, ctypes.c_int, ctypes.c_size_t, _vp],
        "hipDeviceSynchronize": [],
        "hipSetDevice": [ctypes.c_int],
        "hipGetDeviceCount": [ctypes.POINTER(ctypes.c_int)],
        "hipStreamCreate": [ctypes.POINTER(_vp)],
        "hipStreamDestroy": [_vp],
        "hipStreamSynchronize": [_vp],
        "hipEventCreate": [ctypes.POINTER(_vp)],
        "hipEventDestroy": [_vp],
        "hipEventRecord": [_vp, _vp],
        "hipEventSynchronize": [_vp],
        "hipEventElapsedTime": [ctypes.POINTER(ctypes.c_float), _vp, _vp],
        "hipGetErrorString": [ctypes.c_int],
    }
    for name, args in sig.items():
        f = getattr(h, name)
        f.argtypes = args
        f.restype = ctypes.c_int
    h.hipGetErrorString.restype = ctypes.c_char_p
    _hip = h
    return h


def _ok(e, what):
    if e != 0:
        raise RuntimeError(f"{what}: {hip().hipGetErrorString(e).decode()}")


def device_count():
    n = ctypes.c_int(0)
    e = hip().hipGetDeviceCount(ctypes.byref(n))
    return n.value if e == 0 else 0


class DeviceArray:
    """hipMalloc'd buffer with numpy-style shape/dtype bookkeeping."""

    def __init__(self, shape, dtype):
        self.shape = tuple(int(s) for s in np.atleast_1d(shape)) if not isinstance(shape, tuple) else shape
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = _vp()
        _ok(hip().hipMalloc(ctypes.byref(p), max(self.nbytes, 1)), "hipMalloc")
        self.ptr = p.value

    @classmethod
    def from_host(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        _ok(hip().hipMemcpy(d.ptr, a.ctypes.data, a.nbytes, H2D), "hipMemcpy H2D")
        return d

    def to_host(self):
        out = np.empty(self.shape, self.dtype)
        _ok(hip().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, D2H), "hipMemcpy D2H")
        return out

    def free(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def synchronize():
    _ok(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")


class Stream:
    def __init__(self):
        s = _vp()
        _ok(hip().hipStreamCreate(ctypes.byref(s)), "hipStreamCreate")
        self.handle = s.value

    def synchronize(self):
        _ok(hip().hipStreamSynchronize(self.handle), "hipStreamSynchronize")

    def destroy(self):
        if self.handle:
            hip().hipStreamDestroy(self.handle)
            self.handle = None


class Event:
    """hipEvent recorded on a given stream (the stream the kernels run on)."""

    def __init__(self):
        e = _vp()
        _ok(hip().hipEventCreate(ctypes.byref(e)), "hipEventCreate")
        self.handle = e.value

    def record(self, stream=None):
        _ok(hip().hipEventRecord(self.handle, stream.handle if isinstance(stream, Stream) else stream),
            "hipEventRecord")

    def synchronize(self):
        _ok(hip().hipEventSynchronize(self.handle), "hipEventSynchronize")

    def elapsed_ms(self, end):
        t = ctypes.c_float(0)
        _ok(hip().hipEventElapsedTime(ctypes.byref(t), self.handle, end.handle), "hipEventElapsedTime")
        return float(t.value)

    def destroy(self):
        if self.handle:
            hip().hipEventDestroy(self.handle)
            self.handle = None


def gather_points(ptr, shape, idx):
    """Field values at grid indices idx[n] = (iz, iy, ix) of a complex64
    (nz, ny, nx, 3) field in device memory at ptr: (n, 3) complex64, one
    24-byte copy per point (for accuracy samples of fields too large to copy)."""
    nz, ny, nx, _ = shape
    res = np.empty((len(idx), 3), np.complex64)
    h = hip()
    for i, (iz, iy, ix) in enumerate(idx):
        if not (0 <= iz < nz and 0 <= iy < ny and 0 <= ix < nx):
            raise IndexError(f"point {i}: index {(iz, iy, ix)} outside {shape[:3]}")
        off = ((int(iz) * ny + int(iy)) * nx + int(ix)) * 24
        _ok(h.hipMemcpy(res[i].ctypes.data, ptr + off, 24, D2H), "hipMemcpy D2H")
    return res
