"""Coupled-wave structure data model and unpacking (host side).

Mirrors CoupledWaveStructure<double> (CoupledWaveStructure.h:57-148) and the
module tables of cw_allocate / cw_unpack (cpuEvaluator.cpp:34-102).  Waves are
numpy complex128 arrays of shape (n, 6): E0x, E0y, E0z, kx, ky, kz; viewed as
float64 (n, 12) they are .cw records.  .cw files are read and written by the
C ABI (scw_cw_load / scw_cw_save), which validates them.
"""
import ctypes
from dataclasses import dataclass, field
from typing import List

import numpy as np

from . import _native


def _waves(a):
    a = np.ascontiguousarray(a, dtype=np.complex128)
    if a.size == 0:
        return np.zeros((0, 6), np.complex128)
    if a.ndim != 2 or a.shape[1] != 6:
        raise ValueError("waves must have shape (n, 6): E0x, E0y, E0z, kx, ky, kz")
    return a


@dataclass
class HomogeneousLayer:
    """One boundary: z, waves reflected off it (Pr, along -z) and transmitted through it (Pt)."""
    z: float
    Pr: np.ndarray = field(default_factory=lambda: np.zeros((0, 6), np.complex128))
    Pt: np.ndarray = field(default_factory=lambda: np.zeros((0, 6), np.complex128))

    def __post_init__(self):
        self.z = float(self.z)
        self.Pr = _waves(self.Pr)
        self.Pt = _waves(self.Pt)


@dataclass
class WaveTables:
    """cw_allocate / cw_unpack result: W, z_layers, waves_begin, waves_end (FP64)."""
    W: np.ndarray
    z_layers: np.ndarray
    begin: np.ndarray
    end: np.ndarray

    def __post_init__(self):
        self.W = _waves(self.W)
        self.z_layers = np.ascontiguousarray(self.z_layers, np.float64).reshape(-1)
        self.begin = np.ascontiguousarray(self.begin, np.int32).reshape(-1)
        self.end = np.ascontiguousarray(self.end, np.int32).reshape(-1)

    @property
    def layers(self):
        return int(self.z_layers.shape[0])

    def region(self, r):
        return self.W[self.begin[r]:self.end[r]]

    def records(self):
        """float64 (n, 12) view in .cw record order."""
        return np.ascontiguousarray(self.W).view(np.float64).reshape(-1, 12)


class CoupledWaveStructure:
    """Pi (incident waves) and Layers (boundaries with their Pr / Pt waves)."""

    def __init__(self, Pi=None, Layers: List[HomogeneousLayer] = None):
        self.Pi = _waves(Pi if Pi is not None else np.zeros((0, 6)))
        self.Layers = list(Layers or [])

    def n_waves(self):
        return self.Pi.shape[0] + sum(L.Pr.shape[0] + L.Pt.shape[0] for L in self.Layers)

    def unpack(self) -> WaveTables:
        """cw_allocate + cw_unpack order (cpuEvaluator.cpp:34-102): Pi, then Pr_l, Pt_l per layer."""
        L = len(self.Layers)
        parts = [self.Pi]
        begin = np.zeros(L + 1, np.int32)
        end = np.zeros(L + 1, np.int32)
        total = self.Pi.shape[0]
        for l, lay in enumerate(self.Layers):
            total += lay.Pr.shape[0]
            end[l] = total
            begin[l + 1] = total
            total += lay.Pt.shape[0]
            parts += [lay.Pr, lay.Pt]
        end[L] = total
        W = np.concatenate(parts, axis=0) if total else np.zeros((0, 6), np.complex128)
        zl = np.array([lay.z for lay in self.Layers], np.float64)
        return WaveTables(W, zl, begin, end)

    # ---- .cw files through the C ABI ----------------------------------
    def _to_handle(self):
        L = _native.lib()
        h = ctypes.c_void_p()
        _native.check(L.scw_cw_create(ctypes.byref(h)))
        try:
            _native.check(L.scw_cw_set_incident(h, self.Pi.ctypes.data, self.Pi.shape[0]))
            for lay in self.Layers:
                _native.check(L.scw_cw_add_layer(h, lay.z, lay.Pr.ctypes.data, lay.Pr.shape[0], lay.Pt.ctypes.data,
                                                 lay.Pt.shape[0]))
        except Exception:
            L.scw_cw_destroy(h)
            raise
        return h

    def save(self, path):
        L = _native.lib()
        h = self._to_handle()
        try:
            _native.check(L.scw_cw_save(h, str(path).encode()))
        finally:
            L.scw_cw_destroy(h)

    @classmethod
    def load(cls, path):
        L = _native.lib()
        h = ctypes.c_void_p()
        _native.check(L.scw_cw_load(str(path).encode(), ctypes.byref(h)))
        try:
            t = unpack_handle(h)
            n_i = ctypes.c_size_t()
            _native.check(L.scw_cw_counts(h, ctypes.byref(n_i), None, None))
            layers = []
            for l in range(t.layers):
                nr, nt = ctypes.c_size_t(), ctypes.c_size_t()
                _native.check(L.scw_cw_layer(h, l, None, ctypes.byref(nr), ctypes.byref(nt)))
                pr_end = int(t.end[l])
                layers.append(HomogeneousLayer(t.z_layers[l], t.W[pr_end - nr.value:pr_end],
                                               t.W[t.begin[l + 1]:t.begin[l + 1] + nt.value]))
            return cls(t.W[:n_i.value], layers)
        finally:
            L.scw_cw_destroy(h)


def unpack_handle(h) -> WaveTables:
    """scw_cw_unpack of a native .cw handle."""
    L = _native.lib()
    n_l, n_w = ctypes.c_size_t(), ctypes.c_size_t()
    _native.check(L.scw_cw_counts(h, None, ctypes.byref(n_l), ctypes.byref(n_w)))
    W = np.zeros((n_w.value, 6), np.complex128)
    zl = np.zeros(n_l.value, np.float64)
    b = np.zeros(n_l.value + 1, np.int32)
    e = np.zeros(n_l.value + 1, np.int32)
    _native.check(L.scw_cw_unpack(h, W.ctypes.data, zl.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                  b.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                  e.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))))
    return WaveTables(W, zl, b, e)


def load_tables(path) -> WaveTables:
    """Load a .cw file and unpack it with the native loader."""
    L = _native.lib()
    h = ctypes.c_void_p()
    _native.check(L.scw_cw_load(str(path).encode(), ctypes.byref(h)))
    try:
        return unpack_handle(h)
    finally:
        L.scw_cw_destroy(h)
