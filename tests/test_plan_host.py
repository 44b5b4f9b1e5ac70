"""The planner's host half (scw_test_plan_host: scw_plan_create's validation
and FP64 separable decomposition, without a device), on the CPU:

* a focused beam's pupil on an interface: every region is a separable
  kx x ky lattice with M distinct kx and ky (M^2 cells), at most two terms
  per cell (incident and reflected), and the transmitted region one;
* an eigenmode slab (sCWatter's internal-field layout): the slab region is
  dense, with every cell holding every kz (width = mc = its mode count);
* a pupil whose waves have their own kx and ky (no lattice) is not
  separable;
* invalid tables fail as SCW_ERR_ARG before any planning."""
import ctypes

import numpy as np
import pytest

from scwatter_amd import _native, configs
from scwatter_amd.cw import WaveTables


def _plan_host(t: WaveTables):
    L = _native.lib()
    W = np.ascontiguousarray(t.W)
    n = W.shape[0]
    rec = np.zeros((n, 12), np.float64)  # E0 (re, im) x 3, then k (re, im) x 3
    rec[:, 0:6:2], rec[:, 1:6:2] = W[:, :3].real, W[:, :3].imag
    rec[:, 6:12:2], rec[:, 7:12:2] = W[:, 3:].real, W[:, 3:].imag
    zl = np.ascontiguousarray(t.z_layers, np.float64)
    b = np.ascontiguousarray(t.begin, np.int32)
    e = np.ascontiguousarray(t.end, np.int32)
    info = np.zeros((len(zl) + 1, 8), np.int64)
    st = L.scw_test_plan_host(rec.ctypes.data, n, zl.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), len(zl),
                              b.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                              e.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                              info.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)))
    return st, [dict(zip(("separable", "dense", "mx", "my", "mc", "width", "n_terms", "waves"), r)) for r in info]


@pytest.mark.parametrize("M", [5, 21, 81])
def test_pupil_regions_are_lattices(M):
    t = configs.layered_structure([1.0, 1.5 + 0.01j], [0.0], 0.8, 0.6, M, ("x",)).unpack()
    st, info = _plan_host(t)
    assert st == 0
    for r, d in enumerate(info):
        assert d["separable"] and d["mx"] == M and d["my"] == M, (r, d)
        assert d["waves"] == (2 if r == 0 else 1) * M * M
        assert d["width"] == (2 if r == 0 else 1) and d["n_terms"] == d["waves"]


def test_eigenmode_slab_is_dense():
    t = configs.tile_sample_structure(M=5, Z=1, size=(12.0, 12.0, 2.0), lam=0.7, seed=3).unpack()
    st, info = _plan_host(t)
    assert st == 0
    slab = info[1]
    assert slab["separable"] and slab["dense"] and slab["mx"] == 5 and slab["my"] == 5
    assert slab["width"] == slab["mc"] and slab["n_terms"] == slab["mx"] * slab["my"] * slab["mc"]


def test_scattered_pupil_is_not_separable():
    rng = np.random.default_rng(5)
    n = 600
    W = np.column_stack([np.ones(n), np.zeros(n), np.zeros(n), rng.uniform(-1, 1, n), rng.uniform(-1, 1, n),
                         np.ones(n)]).astype(np.complex128)
    st, info = _plan_host(WaveTables(W, np.zeros(0), np.array([0], np.int32), np.array([n], np.int32)))
    assert st == 0 and not info[0]["separable"]


def test_invalid_tables_are_refused():
    t = configs.layered_structure([1.0, 1.5], [0.0], 0.8, 0.6, 5, ("x",)).unpack()
    bad = t.W.copy()
    bad[3, 1] = np.nan
    assert _plan_host(WaveTables(bad, t.z_layers, t.begin, t.end))[0] == -1
    end = np.array(t.end, np.int32)
    end[-1] += 1  # past the table
    assert _plan_host(WaveTables(t.W, t.z_layers, t.begin, end))[0] == -1
