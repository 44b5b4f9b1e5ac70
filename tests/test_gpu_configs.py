"""GPU parity on the benchmark structures at their full size and grid.

Each workload is evaluated on the device through the C ABI; 4096 seeded grid
points are gathered from the device field and compared with the FP64 oracle
(oracle_eval_points_f64, a restatement of cpuEvaluator.cpp:294-334) at the same
coordinates.  Bounds (relative L2 over the sample and all three components):

  GEMM FP32      1e-5
  GEMM BF16 x3   1e-4   (the split form's bar)
  GEMM FP16      2e-3   (reported precision form)
  GEMM BF16      1e-2   (reported precision form)
  direct FP32    2e-5, and 1e-4 on the tile sample, whose 777,924-wave
                 regions are FP32 sums of that many terms

The 41 x 41 order lattices (the 41^2 x 2-polarisation pupil and the 41^2
batch instance) have 41 distinct kx and ky per region, which is what routes
the 16-bit main GEMM to K = 96 with its X fragments staged in LDS over three
K steps; the tests assert those counts from the plan so they stay on that
path.  The batch test takes one instance of each order class of the 64-slab
sweep, with the sweep's own wavelength, angle and seed, on the sweep's
256 x 256 x 64 grid, so the 16-bit mode contraction runs its 64-z workgroup
kernel."""
import os
import time

import numpy as np
import pytest

import oracle
from scwatter_amd import BatchExec, Exec, Plan, autotune, configs, hip

# the fixtures build full-size structures and their oracle samples (up to a
# minute on the host cores), which counts towards the first test of each
pytestmark = [pytest.mark.gpu, pytest.mark.timeout(600)]

NPTS = 4096
TOL = {"fp32": 1e-5, "bf16x3": 1e-4, "fp16": 2e-3, "bf16": 1e-2}
NTHR = min(16, os.cpu_count() or 1)


def _log(msg):
    print(msg, flush=True)


def _sample(shape, n, seed):
    rng = np.random.default_rng(seed)
    nz, ny, nx = shape[:3]
    return np.stack([rng.integers(0, nz, n), rng.integers(0, ny, n), rng.integers(0, nx, n)], 1)


def _check(E, ref, tol, what):
    rl2 = oracle.rel_l2(E, ref)
    mx = oracle.max_abs(E, ref) / max(np.abs(ref).max(), 1e-30)
    _log(f"{what}: rel-L2 {rl2:.3e} max-abs/max {mx:.3e} ({len(E)} points)")
    assert np.isfinite(E).all(), what
    assert rl2 <= tol, f"{what}: rel-L2 {rl2:.3e} > {tol:.1e}"


class Workload:
    def __init__(self, name, tables, xs, ys, zs, seed):
        t0 = time.perf_counter()
        self.name, self.t, self.xs, self.ys, self.zs = name, tables, xs, ys, zs
        self.shape = (zs.size, ys.size, xs.size, 3)
        self.idx = _sample(self.shape, NPTS, seed)
        pts = np.stack([xs[self.idx[:, 2]], ys[self.idx[:, 1]], zs[self.idx[:, 0]]], 1)
        self.ref = oracle.eval_points_f64(tables.W, tables.z_layers, tables.begin, tables.end, pts, nthreads=NTHR)
        self.plan = Plan(tables)
        self.out = hip.DeviceArray(self.shape, np.complex64)
        _log(f"{name}: {tables.W.shape[0]} waves, grid {xs.size}x{ys.size}x{zs.size}, oracle at {NPTS} points "
             f"({time.perf_counter() - t0:.1f} s)")

    def run(self, form, dtype, tuned=None):
        """Evaluate in `form`/`dtype`; with `tuned` (a workload key of the
        committed tuning table), under that table's entry for this exec: the
        launch shape, z-run tiles and form (stage by stage or single launch)
        that the bench runs."""
        ex = Exec(self.plan, self.xs, self.ys, self.zs, form=form, dtype=dtype)
        assert ex.info()["form"] == form
        if tuned:
            entry = autotune.apply_cached(ex, autotune.shape_key(ex, tuned))
            assert entry is not None, f"no tuning-table entry for {tuned} in {dtype}"
            shown = {k: entry.get(k) for k in ("coef_items", "main_blocks", "modal_run", "coef_run", "single")}
            _log(f"{self.name} {dtype}: tuned entry {shown}")
        ex.run(self.out.ptr)
        hip.synchronize()
        E = hip.gather_points(self.out.ptr, self.shape, self.idx)
        ex.destroy()
        return E


@pytest.fixture(scope="module")
def c2():
    cw, (xs, ys, zs) = configs.config2()
    t = cw.unpack()
    del cw
    return Workload("tile sample 21^2 orders, 20 sublayers", t, xs, ys, zs, 2)


@pytest.fixture(scope="module")
def c3():
    cw, (xs, ys, zs) = configs.config3()
    return Workload("41^2 pupil x 2 polarisations", cw.unpack(), xs, ys, zs, 3)


@pytest.fixture(scope="module")
def c5():
    cw, (xs, ys, zs) = configs.config5()
    return Workload("16-layer stack, 31^2 pupil", cw.unpack(), xs, ys, zs, 5)


@pytest.mark.parametrize("form,dtype", [("gemm", "bf16x3"), ("gemm", "fp32"), ("gemm", "bf16"), ("gemm", "fp16"),
                                        ("direct", "fp32")])
def test_tile_sample_full(c2, form, dtype):
    info = c2.plan.region_info(11)
    assert info["n_waves"] == 4 * 441 * 441 and info["mx"] == info["my"] == 21 and info["width"] == 4 * 441
    _check(c2.run(form, dtype), c2.ref, 1e-4 if form == "direct" else TOL[dtype], f"c2 {form} {dtype}")


@pytest.mark.parametrize("form,dtype", [("gemm", "bf16x3"), ("gemm", "fp32"), ("gemm", "bf16"), ("gemm", "fp16"),
                                        ("direct", "fp32")])
def test_pupil_41_two_polarisations_full(c3, form, dtype):
    for r in (0, 1):
        info = c3.plan.region_info(r)
        assert info["separable"] and info["mx"] == info["my"] == 41  # 33..48: K = 96, X staged over 3 steps
    _check(c3.run(form, dtype), c3.ref, 2e-5 if form == "direct" else TOL[dtype], f"c3 {form} {dtype}")


@pytest.mark.parametrize("form,dtype", [("gemm", "bf16x3"), ("gemm", "fp32"), ("gemm", "fp16"), ("direct", "fp32")])
def test_layer_stack_16_full(c5, form, dtype):
    assert c5.plan.layers == 17
    _check(c5.run(form, dtype), c5.ref, 2e-5 if form == "direct" else TOL[dtype], f"c5 {form} {dtype}")


@pytest.mark.parametrize("workload,dtype", [(w, d) for w in ("c2", "c3") for d in ("bf16x3", "fp32", "bf16", "fp16")] +
                         [("c5", d) for d in ("bf16x3", "fp32", "fp16")])
def test_full_size_tuned_shapes(request, workload, dtype):
    """The GEMM form under the committed tuning table's entry for each
    full-size workload, as the bench runs it (the 16-layer stack in BF16x3
    takes the single-launch form), on the same oracle sample and bars."""
    w = request.getfixturevalue(workload)
    _check(w.run("gemm", dtype, tuned=workload), w.ref, TOL[dtype], f"{workload} gemm {dtype}, tuned launch shape")


# Sample points per instance of the batch test (4096 in all; fewer on the
# larger instances, whose oracle sums take longer per point).
C4_POINTS = {11: 1536, 21: 1280, 31: 768, 41: 512}


@pytest.fixture(scope="module")
def c4_classes():
    t0 = time.perf_counter()
    params = configs.config4_instances()
    pick = {}
    for p in params:  # the first instance of each order class, as the sweep draws it
        pick.setdefault(p["M"], p)
    assert sorted(pick) == [11, 21, 31, 41]
    xs, ys, zs = configs.config4_grid()
    plans, refs, idxs = [], [], []
    for k, M in enumerate(sorted(pick)):
        t = configs.config4_structure(pick[M]).unpack()
        plans.append(Plan(t))
        n = C4_POINTS[M]
        idx = _sample((zs.size, ys.size, xs.size, 3), n, 40 + M)
        pts = np.stack([xs[idx[:, 2]], ys[idx[:, 1]], zs[idx[:, 0]]], 1)
        refs.append(oracle.eval_points_f64(t.W, t.z_layers, t.begin, t.end, pts, nthreads=NTHR))
        idx = idx.copy()
        idx[:, 0] += k * zs.size  # the batch output stacks the instances along z
        idxs.append(idx)
        _log(f"c4 class {M}^2: {t.W.shape[0]} waves, oracle at {n} points ({time.perf_counter() - t0:.1f} s)")
        del t
    return plans, (xs, ys, zs), np.concatenate(idxs), np.concatenate(refs), [pick[M] for M in sorted(pick)]


@pytest.mark.parametrize("dtype", ["bf16x3", "fp32", "fp16", "bf16"])
def test_batch_order_classes_full(c4_classes, dtype):
    plans, (xs, ys, zs), idx, ref, params = c4_classes
    for plan, p in zip(plans, params):
        info = plan.region_info(1)
        assert info["mx"] == info["my"] == p["M"] and info["width"] == 4 * p["M"] ** 2
    bx = BatchExec(plans, xs, ys, [zs] * len(plans), dtype=dtype)
    out = hip.DeviceArray(bx.shape, np.complex64)
    bx.run(out.ptr)
    hip.synchronize()
    E = hip.gather_points(out.ptr, bx.shape, idx)
    bx.destroy()
    out.free()
    _check(E, ref, TOL[dtype], f"c4 order classes {dtype}")


@pytest.mark.parametrize("dtype", ["fp32", "bf16x3"])
def test_batch_main_launch_per_k_bit_identical(c4_classes, dtype, monkeypatch):
    """The main GEMM as one launch at the batch's largest K (the 16-bit
    default) or as one launch per kernel K over the regrouped z entries (the
    FP32 default; SCW_MAIN_KSPLIT forces either): the same field, bit for bit,
    on the four order classes (kernel K 32 to 128)."""
    plans, (xs, ys, zs), idx, ref, params = c4_classes
    fields = []
    for split in ("0", "1"):
        monkeypatch.setenv("SCW_MAIN_KSPLIT", split)
        bx = BatchExec(plans, xs, ys, [zs] * len(plans), dtype=dtype)
        out = hip.DeviceArray(bx.shape, np.complex64)
        bx.run(out.ptr)
        hip.synchronize()
        fields.append(out.to_host())
        bx.destroy()
        out.free()
    assert np.array_equal(fields[0], fields[1])
    E = fields[1][idx[:, 0], idx[:, 1], idx[:, 2]]
    _check(E, ref, TOL[dtype], f"c4 order classes {dtype}, main GEMM per kernel K")


# Points per instance of the whole-sweep test (64 instances: 4096 in all).
SWEEP_POINTS = 64


@pytest.fixture(scope="module")
def c4_sweep():
    """The whole batched sweep as the bench runs it: all 64 instances of
    config4_instances() (orders drawn from {11, 21, 31, 41}^2) on their
    256 x 256 x 64 grids, with the FP64 oracle at 64 seeded points each."""
    t0 = time.perf_counter()
    params = configs.config4_instances()
    xs, ys, zs = configs.config4_grid()
    rng = np.random.default_rng(4096)
    idx = [np.stack([rng.integers(0, n, SWEEP_POINTS) for n in (zs.size, ys.size, xs.size)], 1) for _ in params]
    oracle._load()

    def build(i, t):
        pts = np.stack([xs[idx[i][:, 2]], ys[idx[i][:, 1]], zs[idx[i][:, 0]]], 1)
        ref = oracle.eval_points_f64(t.W, t.z_layers, t.begin, t.end, pts, nthreads=max(1, NTHR // 8))
        return Plan(t), ref
    built = configs.map_instances(build, params, workers=8)
    plans = [b[0] for b in built]
    ref = np.concatenate([b[1] for b in built])
    gidx = np.concatenate([ix + np.array([i * zs.size, 0, 0]) for i, ix in enumerate(idx)])
    _log(f"c4 sweep: {len(plans)} instances, oracle at {len(gidx)} points ({time.perf_counter() - t0:.1f} s)")
    return plans, (xs, ys, zs), gidx, ref


@pytest.mark.parametrize("shape", ["default", "tuned"])
@pytest.mark.parametrize("dtype", ["bf16x3", "fp32", "fp16", "bf16"])
def test_batch_sweep_64_full(c4_sweep, dtype, shape):
    """All 64 instances in one batch run (6.44 GB field), sampled on the
    device at 4,096 points against the FP64 oracle: with the library's default
    launch shape, and with the committed tuning table's entry for this batch
    (launch shape, mode slices, z-run tiles and form: what the bench runs)."""
    plans, (xs, ys, zs), gidx, ref = c4_sweep
    bx = BatchExec(plans, xs, ys, [zs] * len(plans), dtype=dtype)
    if shape == "tuned":
        entry = autotune.apply_cached(bx, autotune.shape_key(bx, "c4"))
        assert entry is not None, f"no tuning-table entry for the batched sweep in {dtype}"
        shown = {k: entry.get(k) for k in ("coef_items", "main_blocks", "modal_run", "coef_run", "single")}
        _log(f"c4 whole sweep {dtype}: tuned entry {shown}")
    out = hip.DeviceArray(bx.shape, np.complex64)
    bx.run(out.ptr)
    hip.synchronize()
    E = hip.gather_points(out.ptr, bx.shape, gidx)
    bx.destroy()
    out.free()
    _check(E, ref, TOL[dtype], f"c4 whole sweep {dtype}, {shape} launch shape")
