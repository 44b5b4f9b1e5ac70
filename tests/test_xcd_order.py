"""Dispatch order of the 64-z mode contraction's workgroup groups over the XCDs
(scw_test_xcd_order, device-free).

Workgroup b of a launch runs on XCD b mod nxcd.  The order must put every
group whole on one XCD (its workgroups share one run's phase fragments, read
through that XCD's L2), list each XCD's groups longest first, balance the
XCDs' step totals the way longest-processing-time-first does, place every
workgroup exactly once, and pad only where an XCD's queue is shorter than
the longest."""
import ctypes

import numpy as np
import pytest

from scwatter_amd import _native

_i64p = ctypes.POINTER(ctypes.c_int64)


def order(wgs, steps, nxcd):
    L = _native.lib()
    wgs = np.ascontiguousarray(wgs, np.int64)
    steps = np.ascontiguousarray(steps, np.int64)
    cap = nxcd * max(int(wgs.sum()), 1)  # at most one row of slots per workgroup
    out = np.empty(max(cap, 1), np.int64)
    n = ctypes.c_int64()
    _native.check(L.scw_test_xcd_order(len(wgs), wgs.ctypes.data_as(_i64p), steps.ctypes.data_as(_i64p), nxcd,
                                       out.ctypes.data_as(_i64p), cap, ctypes.byref(n)))
    return out[:n.value]


def check(wgs, steps, nxcd):
    o = order(wgs, steps, nxcd)
    first = np.concatenate([[0], np.cumsum(wgs)])
    group_of = np.repeat(np.arange(len(wgs)), wgs)
    assert len(o) % nxcd == 0
    placed = o[o >= 0]
    assert sorted(placed.tolist()) == list(range(int(first[-1])))  # every workgroup once
    xcd = {}
    for slot, w in enumerate(o):
        if w < 0:
            continue
        g = int(group_of[w])
        assert xcd.setdefault(g, slot % nxcd) == slot % nxcd  # a group never spans two XCDs
    load = np.zeros(nxcd, np.int64)
    for x in range(nxcd):
        col = o[x::nxcd]
        live = col[col >= 0]
        assert (col[len(live):] < 0).all()  # padding only after the XCD's queue
        lens = [steps[group_of[w]] for w in live]
        assert all(a >= b for a, b in zip(lens, lens[1:]))  # longest first within the XCD
        load[x] = sum(lens)
    if len(wgs):
        # LPT: the most-loaded XCD carries at most one group's work above the least-loaded
        assert load.max() - load.min() <= max(int(w * s) for w, s in zip(wgs, steps))
    return o


def test_sweep_like_groups():
    # a 64-slab sweep: 16 slabs of each order class, workgroups per (run, slot)
    # group and 16-mode steps per group as the batch's 11^2 .. 41^2 classes
    rng = np.random.default_rng(3)
    classes = [(3, 31), (6, 111), (13, 241), (20, 421)]
    wgs, steps = [], []
    for _ in range(64):
        w, t = classes[rng.integers(4)]
        for slot in range(int(rng.integers(1, 3))):
            wgs.append(w)
            steps.append(t // 2 if slot else t - t // 2)
    check(np.array(wgs), np.array(steps), 8)


@pytest.mark.parametrize("seed", range(5))
@pytest.mark.parametrize("nxcd", [2, 3, 8])
def test_random_groups(seed, nxcd):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(1, 40))
    check(rng.integers(1, 25, n), rng.integers(1, 500, n), nxcd)


def test_fewer_groups_than_xcds():
    o = check(np.array([5, 2]), np.array([10, 30]), 8)
    # the longer group leads XCD 0, the other goes whole to XCD 1
    assert o[0] == 5 and o[1] == 0 and (o[2:8] < 0).all()


def test_empty_and_errors():
    assert len(order(np.array([], np.int64), np.array([], np.int64), 8)) == 0
    L = _native.lib()
    n = ctypes.c_int64()
    one = np.array([4], np.int64)
    out = np.empty(2, np.int64)
    st = L.scw_test_xcd_order(1, one.ctypes.data_as(_i64p), one.ctypes.data_as(_i64p), 8,
                              out.ctypes.data_as(_i64p), 2, ctypes.byref(n))
    assert _native.STATUS[st] == "SCW_ERR_ARG"
    assert n.value == 32  # 4 workgroups on one XCD: 4 rows of 8 slots
