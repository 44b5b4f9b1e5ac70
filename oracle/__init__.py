"""Oracle — test infrastructure only (ctypes wrapper of liboracle.so).

FP64 restatement of scwatter's CPU evaluator (cpuEvaluator.cpp:294-407) and an
FP32 reference-semantics restatement, both in scw_oracle.cpp.  Imported only by
tests/, by __graft_entry__.smoke() and by the CPU-baseline leg of bench.py, as
the checker; the shipped library never uses it.

Wave arrays are complex128 of shape (n, 6): columns E0x, E0y, E0z, kx, ky, kz —
viewed as float64 (n, 12) this is the .cw record order (CoupledWaveStructure.h:7-55).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_ip = ctypes.POINTER(ctypes.c_int32)
_i64 = ctypes.c_int64


def build():
    """Compile liboracle.so with the committed Makefile (g++)."""
    subprocess.run(["make", "-s", "-C", _HERE, "-B", "liboracle.so"], check=True)


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB):
        build()
    lib = ctypes.CDLL(_LIB)
    lib.oracle_region_of.restype = ctypes.c_int
    lib.oracle_region_of.argtypes = [ctypes.c_double, _dp, ctypes.c_int, ctypes.c_double]
    lib.oracle_eval_grid_f64.restype = None
    lib.oracle_eval_grid_f64.argtypes = [_dp, _i64, _dp, ctypes.c_int, _ip, _ip, _dp, _i64, _dp, _i64,
                                         _dp, _i64, ctypes.c_double, _dp, ctypes.c_int]
    lib.oracle_eval_points_f64.restype = None
    lib.oracle_eval_points_f64.argtypes = [_dp, _i64, _dp, ctypes.c_int, _ip, _ip, _dp, _i64,
                                           ctypes.c_double, _dp, ctypes.c_int]
    lib.refcpu_unpack_f32.restype = None
    lib.refcpu_unpack_f32.argtypes = [_dp, _i64, _fp]
    for name in ("refcpu_eval_xy", "refcpu_eval_yz", "refcpu_eval_xz"):
        f = getattr(lib, name)
        f.restype = None
        f.argtypes = [_fp, _fp, ctypes.c_int, _ip, _ip, _fp, ctypes.c_float, ctypes.c_float,
                      ctypes.c_float, ctypes.c_float, _i64]
    lib.refcpu_eval_grid.restype = None
    lib.refcpu_eval_grid.argtypes = [_fp, _fp, ctypes.c_int, _ip, _ip, _fp, _i64, _fp, _i64, _fp, _i64,
                                     _fp, ctypes.c_int]
    _lib = lib
    return lib


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


def _ptr(a, ct):
    return a.ctypes.data_as(ctypes.POINTER(ct))


def _tables(W, z_layers, begin, end):
    W = _c(W, np.complex128)
    assert W.ndim == 2 and W.shape[1] == 6
    zl = _c(z_layers, np.float64)
    b = _c(begin, np.int32)
    e = _c(end, np.int32)
    assert b.shape == e.shape == (zl.shape[0] + 1,)
    assert (b >= 0).all() and (e <= W.shape[0]).all() and (b <= e).all()
    return W, zl, b, e


def region_of(z, z_layers, tol=0.0):
    zl = _c(z_layers, np.float64)
    return _load().oracle_region_of(float(z), _ptr(zl, ctypes.c_double), zl.shape[0], float(tol))


def eval_grid_f64(W, z_layers, begin, end, xs, ys, zs, tol=0.0, nthreads=0):
    """FP64 field on the grid xs x ys x zs -> complex128 (nz, ny, nx, 3)."""
    W, zl, b, e = _tables(W, z_layers, begin, end)
    xs, ys, zs = (_c(v, np.float64) for v in (xs, ys, zs))
    out = np.empty((zs.size, ys.size, xs.size, 3), np.complex128)
    _load().oracle_eval_grid_f64(_ptr(W.view(np.float64), ctypes.c_double), W.shape[0],
                                 _ptr(zl, ctypes.c_double), zl.shape[0], _ptr(b, ctypes.c_int32),
                                 _ptr(e, ctypes.c_int32), _ptr(xs, ctypes.c_double), xs.size,
                                 _ptr(ys, ctypes.c_double), ys.size, _ptr(zs, ctypes.c_double), zs.size,
                                 float(tol), _ptr(out.view(np.float64), ctypes.c_double), int(nthreads))
    return out


def eval_points_f64(W, z_layers, begin, end, pts, tol=0.0, nthreads=0):
    """FP64 field at points (n, 3) -> complex128 (n, 3)."""
    W, zl, b, e = _tables(W, z_layers, begin, end)
    pts = _c(pts, np.float64).reshape(-1, 3)
    out = np.empty((pts.shape[0], 3), np.complex128)
    _load().oracle_eval_points_f64(_ptr(W.view(np.float64), ctypes.c_double), W.shape[0],
                                   _ptr(zl, ctypes.c_double), zl.shape[0], _ptr(b, ctypes.c_int32),
                                   _ptr(e, ctypes.c_int32), _ptr(pts, ctypes.c_double), pts.shape[0],
                                   float(tol), _ptr(out.view(np.float64), ctypes.c_double), int(nthreads))
    return out


def unpack_f32(W):
    """cw_unpack narrowing: each re/im part rounded to float -> complex64 (n, 6)."""
    W = _c(W, np.complex128)
    out = np.empty(W.shape, np.complex64)
    _load().refcpu_unpack_f32(_ptr(W.view(np.float64), ctypes.c_double), W.shape[0],
                              _ptr(out.view(np.float32), ctypes.c_float))
    return out


def _ftables(W, z_layers, begin, end):
    W, zl, b, e = _tables(W, z_layers, begin, end)
    Wf = unpack_f32(W)
    zf = _c(zl.astype(np.float32), np.float32)
    return Wf, zf, b, e


def refcpu_plane(kind, W, z_layers, begin, end, a_start, b_start, fixed, d, N):
    """cpu_cw_evaluate_{xy,yz,xz} with the reference's FP32 semantics -> complex64 (N, N, 3)."""
    Wf, zf, b, e = _ftables(W, z_layers, begin, end)
    out = np.empty((N, N, 3), np.complex64)
    f = getattr(_load(), "refcpu_eval_" + kind)
    f(_ptr(Wf.view(np.float32), ctypes.c_float), _ptr(zf, ctypes.c_float), zf.shape[0],
      _ptr(b, ctypes.c_int32), _ptr(e, ctypes.c_int32), _ptr(out.view(np.float32), ctypes.c_float),
      float(a_start), float(b_start), float(fixed), float(d), int(N))
    return out


def refcpu_grid(W, z_layers, begin, end, xs, ys, zs, nthreads=1):
    """Reference-semantics FP32 evaluation on a grid of float coordinates -> complex64 (nz, ny, nx, 3)."""
    Wf, zf, b, e = _ftables(W, z_layers, begin, end)
    xs, ys, zs = (_c(v, np.float32) for v in (xs, ys, zs))
    out = np.empty((zs.size, ys.size, xs.size, 3), np.complex64)
    _load().refcpu_eval_grid(_ptr(Wf.view(np.float32), ctypes.c_float), _ptr(zf, ctypes.c_float), zf.shape[0],
                             _ptr(b, ctypes.c_int32), _ptr(e, ctypes.c_int32), _ptr(xs, ctypes.c_float),
                             xs.size, _ptr(ys, ctypes.c_float), ys.size, _ptr(zs, ctypes.c_float), zs.size,
                             _ptr(out.view(np.float32), ctypes.c_float), int(nthreads))
    return out


def rel_l2(a, ref):
    a = np.asarray(a, np.complex128)
    ref = np.asarray(ref, np.complex128)
    den = np.linalg.norm(ref.ravel())
    return float(np.linalg.norm((a - ref).ravel()) / (den if den > 0 else 1.0))


def max_abs(a, ref):
    return float(np.max(np.abs(np.asarray(a, np.complex128) - np.asarray(ref, np.complex128)))) if np.size(a) else 0.0
