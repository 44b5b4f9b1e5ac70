"""GPU parity on physical eigen-tables: the tile-sample workloads with their
eigenvalues and mode amplitudes from the host sample solver (sCWatter's Phi /
eigen-solve / boundary pipeline, scwatter_amd/sample_solver.py) instead of
seeded synthetic ones.  Bounds as for the synthetic tables (relative L2 over
4096 seeded points against the FP64 oracle): GEMM FP32 1e-5, BF16 x3 1e-4,
FP16 2e-3, BF16 1e-2."""
import os
import time

import numpy as np
import pytest

import oracle
from scwatter_amd import BatchExec, Exec, Plan, configs, hip

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(900)]
TOL = {"fp32": 1e-5, "bf16x3": 1e-4, "fp16": 2e-3, "bf16": 1e-2}
NTHR = min(16, os.cpu_count() or 1)


def _log(msg):
    print(msg, flush=True)


def _sample(shape, n, seed):
    rng = np.random.default_rng(seed)
    return np.stack([rng.integers(0, shape[0], n), rng.integers(0, shape[1], n), rng.integers(0, shape[2], n)], 1)


def _check(E, ref, tol, what):
    rl2 = oracle.rel_l2(E, ref)
    mx = oracle.max_abs(E, ref) / max(np.abs(ref).max(), 1e-30)
    _log(f"{what}: rel-L2 {rl2:.3e} max-abs/max {mx:.3e} ({len(E)} points)")
    assert np.isfinite(E).all() and rl2 <= tol, f"{what}: rel-L2 {rl2:.3e} > {tol:.1e}"


@pytest.fixture(scope="module")
def c2p():
    t0 = time.perf_counter()
    cw, (xs, ys, zs) = configs.config2_physical()
    t = cw.unpack()
    del cw
    idx = _sample((zs.size, ys.size, xs.size), 4096, 22)
    pts = np.stack([xs[idx[:, 2]], ys[idx[:, 1]], zs[idx[:, 0]]], 1)
    ref = oracle.eval_points_f64(t.W, t.z_layers, t.begin, t.end, pts, nthreads=NTHR)
    _log(f"physical tile sample: {t.W.shape[0]} waves, solved and sampled in {time.perf_counter() - t0:.1f} s")
    return Plan(t), (xs, ys, zs), idx, ref


@pytest.mark.parametrize("dtype", ["bf16x3", "fp32", "fp16", "bf16"])
def test_physical_tile_sample_full(c2p, dtype):
    plan, (xs, ys, zs), idx, ref = c2p
    info = plan.region_info(11)
    assert info["separable"] and info["width"] == 4 * 441
    ex = Exec(plan, xs, ys, zs, form="gemm", dtype=dtype)
    out = hip.DeviceArray(ex.shape, np.complex64)
    ex.run(out.ptr)
    hip.synchronize()
    _check(hip.gather_points(out.ptr, ex.shape, idx), ref, TOL[dtype], f"physical c2 gemm {dtype}")


@pytest.fixture(scope="module")
def c4p():
    t0 = time.perf_counter()
    pick = {}
    for p in configs.config4_instances():
        pick.setdefault(p["M"], p)
    params = [pick[11], pick[21]]
    xs, ys, zs = configs.config4_grid()
    plans, refs, idxs = [], [], []
    for k, p in enumerate(params):
        t = configs.config4_physical_structure(p).unpack()
        plans.append(Plan(t))
        idx = _sample((zs.size, ys.size, xs.size), 2048, 60 + p["M"])
        pts = np.stack([xs[idx[:, 2]], ys[idx[:, 1]], zs[idx[:, 0]]], 1)
        refs.append(oracle.eval_points_f64(t.W, t.z_layers, t.begin, t.end, pts, nthreads=NTHR))
        idx = idx.copy()
        idx[:, 0] += k * zs.size
        idxs.append(idx)
    _log(f"physical sweep instances {[(p['M'], round(p['lam'], 3), round(p['angle'], 1)) for p in params]} "
         f"({time.perf_counter() - t0:.1f} s)")
    return plans, (xs, ys, zs), np.concatenate(idxs), np.concatenate(refs)


@pytest.mark.parametrize("dtype", ["bf16x3", "fp32", "fp16", "bf16"])
def test_physical_batch_instances(c4p, dtype):
    plans, (xs, ys, zs), idx, ref = c4p
    bx = BatchExec(plans, xs, ys, [zs] * len(plans), dtype=dtype)
    out = hip.DeviceArray(bx.shape, np.complex64)
    bx.run(out.ptr)
    hip.synchronize()
    _check(hip.gather_points(out.ptr, bx.shape, idx), ref, TOL[dtype], f"physical sweep instances gemm {dtype}")
