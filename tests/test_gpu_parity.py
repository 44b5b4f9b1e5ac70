"""GPU parity: the HIP kernels, called through the C ABI, against the FP64
oracle and the mpmath golden vectors.

Tolerances (relative L2 over all three components, and max-abs relative to
the field's max):
  direct FP32          rel-L2 <= 1e-5   (config-1 class fields, |k.r| <= ~150)
  GEMM FP32, BF16 x3   rel-L2 <= 1e-5 / 1e-4
  GEMM BF16, FP16      rel-L2 <= 1e-2 (reported precision forms, not the bar)
"""
import os

import numpy as np
import pytest

import oracle
from scwatter_amd import Exec, Plan, configs, cpu_cw_evaluate, gpu_cw_evaluate
from scwatter_amd.cw import CoupledWaveStructure

pytestmark = pytest.mark.gpu

TOL = {"fp32": 1e-5, "bf16x3": 1e-4, "bf16": 1e-2, "fp16": 1e-2}


def _log(msg):
    print(msg, flush=True)


def _check(E, ref, tol, what):
    rl2 = oracle.rel_l2(E, ref)
    mx = oracle.max_abs(E, ref) / max(np.abs(ref).max(), 1e-30)
    _log(f"{what}: rel-L2 {rl2:.3e} max-abs/max {mx:.3e}")
    assert np.isfinite(E).all(), what
    assert rl2 <= tol, f"{what}: rel-L2 {rl2:.3e} > {tol:.1e}"
    return rl2


@pytest.mark.parametrize("case", ["A", "B", "C", "D"])
def test_direct_matches_mpmath(golden, case):
    g = golden[case]
    _log(f"golden case {case}")
    from scwatter_amd.cw import WaveTables
    plan = Plan(WaveTables(g["W"], g["zl"], g["begin"], g["end"]))
    ex = Exec(plan, g["xs"], g["ys"], g["zs"], float(g["tol"]), form="direct")
    E = ex.run_host()
    _check(E, g["E"], 1e-5, f"direct golden {case}")


def test_direct_matches_mpmath_points(golden):
    g = golden["E"]
    from scwatter_amd.cw import WaveTables
    plan = Plan(WaveTables(g["W"], g["zl"], g["begin"], g["end"]))
    out = []
    for p in g["pts"]:
        out.append(Exec(plan, [p[0]], [p[1]], [p[2]], 0.0, form="direct").run_host().reshape(3))
    _check(np.array(out), g["E"], 1e-5, "direct golden E (points)")


@pytest.fixture(scope="module")
def c1():
    cw, (xs, ys, zs) = configs.config1()
    t = cw.unpack()
    ref = oracle.eval_grid_f64(t.W, t.z_layers, t.begin, t.end, xs, ys, zs)
    return Plan(t), (xs, ys, zs), ref


def test_c1_direct_full_grid(c1):
    plan, g, ref = c1
    ex = Exec(plan, *g, form="direct")
    assert ex.info()["form"] == "direct"
    _check(ex.run_host(), ref, 1e-5, "C1 direct 64^3")


@pytest.mark.parametrize("dtype", ["fp32", "bf16x3", "bf16", "fp16"])
def test_c1_gemm_full_grid(c1, dtype):
    plan, g, ref = c1
    for r in range(2):
        assert plan.region_info(r)["separable"]
    ex = Exec(plan, *g, form="gemm", dtype=dtype)
    info = ex.info()
    assert info["form"] == "gemm" and info["dtype"] == dtype
    _check(ex.run_host(), ref, TOL[dtype], f"C1 gemm {dtype} 64^3")


def test_c1_auto_picks_gemm(c1):
    plan, g, _ = c1
    assert Exec(plan, *g).info()["form"] == "gemm"


def test_c1_units(c1):
    plan, (xs, ys, zs), _ = c1
    info = Exec(plan, xs, ys, zs, form="direct").info()
    n_below = int((zs < 0).sum())
    assert info["units"] == 64 * 64 * (n_below * 882 + (64 - n_below) * 441)


@pytest.fixture(scope="module")
def tile_small():
    cw = configs.tile_sample_structure(M=5, Z=4, size=(16.0, 16.0, 4.0), lam=1.0, seed=7)
    t = cw.unpack()
    xs, ys, zs = configs.axis(-8, 8, 24), configs.axis(-8, 8, 20), configs.axis(-3, 3, 26)
    ref = oracle.eval_grid_f64(t.W, t.z_layers, t.begin, t.end, xs, ys, zs)
    return Plan(t), (xs, ys, zs), ref


def test_tile_sample_direct(tile_small):
    plan, g, ref = tile_small
    _check(Exec(plan, *g, form="direct").run_host(), ref, 1e-5, "tile sample direct")


@pytest.mark.parametrize("dtype", ["fp32", "bf16x3"])
def test_tile_sample_gemm(tile_small, dtype):
    plan, g, ref = tile_small
    info = plan.region_info(2)
    assert info["separable"] and info["n_waves"] == 4 * 25 * 25 and info["mc"] == 100 and info["width"] == 100
    _check(Exec(plan, *g, form="gemm", dtype=dtype).run_host(), ref, TOL[dtype], f"tile sample gemm {dtype}")


@pytest.mark.parametrize("dtype", ["fp32", "bf16x3", "bf16", "fp16"])
@pytest.mark.parametrize("nx,ny", [(1, 33), (3, 17), (45, 5), (50, 19), (67, 1)])
def test_gemm_ragged_grids(tile_small, dtype, nx, ny):
    """Grids whose rows are not whole 16-byte groups (odd nx: 6 nx floats per
    row), whose last 256-column block is partial (nx = 50, 67) and whose y
    extent is not a whole number of 16-row tiles."""
    plan, _, _ = tile_small
    xs, ys, zs = configs.axis(-7, 7.5, nx), configs.axis(-6, 8, ny), configs.axis(-3, 3, 9)
    t = plan.tables
    ref = oracle.eval_grid_f64(t.W, t.z_layers, t.begin, t.end, xs, ys, zs)
    _check(Exec(plan, xs, ys, zs, form="gemm", dtype=dtype).run_host(), ref, TOL[dtype],
           f"ragged grid {nx}x{ny}x{zs.size} gemm {dtype}")


@pytest.fixture(scope="module")
def slab_dense():
    cw = configs.tile_sample_structure(M=7, Z=1, size=(12.0, 12.0, 2.0), lam=0.7, seed=3, incidence_deg=12.0)
    t = cw.unpack()
    # 75 z inside the slab: one stretch of the dense region, split into runs
    # of 38 and 37 z (5 row tiles, in the 8-tile class of the MFMA mode
    # contraction, with partly filled tiles); 5 z outside it
    zs = np.concatenate([configs.axis(-0.95, 0.95, 75), configs.axis(1.05, 1.6, 5)])
    xs, ys = configs.axis(-6, 6, 18), configs.axis(-6, 6, 14)
    ref = oracle.eval_grid_f64(t.W, t.z_layers, t.begin, t.end, xs, ys, zs)
    return Plan(t), (xs, ys, zs), ref


@pytest.mark.parametrize("dtype", ["fp32", "bf16x3", "bf16", "fp16"])
def test_dense_slab_mode_contraction(slab_dense, dtype, monkeypatch):
    """Long z stretches through a dense eigenmode region: the MFMA mode
    contraction (runs of up to 64 z) and the VALU term contraction (forced
    with SCW_MODAL_MIN_RUN) both match the oracle."""
    plan, g, ref = slab_dense
    info = plan.region_info(1)  # every cell holds every kz: dense
    assert info["width"] == info["mc"] and info["n_terms"] == info["mx"] * info["my"] * info["mc"]
    _check(Exec(plan, *g, form="gemm", dtype=dtype).run_host(), ref, TOL[dtype], f"dense slab modal {dtype}")
    monkeypatch.setenv("SCW_MODAL_MIN_RUN", "1000000")
    _check(Exec(plan, *g, form="gemm", dtype=dtype).run_host(), ref, TOL[dtype], f"dense slab VALU {dtype}")


@pytest.mark.parametrize("dtype,tol", [("fp16", 2e-3), ("bf16x3", 1e-4)])
def test_strongly_evanescent_slab(dtype, tol):
    """A 4 um slab at 0.4 um with evanescent modes up to Im(kz) = 6.3 / um:
    across the slab exp(i kz (z - zref)) spans e^(+-12.6), past FP16's range,
    and the amplitudes re-referenced to the slab's middle carry the inverse
    factor.  The FP16 mode contraction moves a power of two per mode from the
    phase factor into the amplitudes; without it FP16 lost the weak factors
    (rel-L2 about 7e-3 on the 64-slab batch, worse than BF16)."""
    cw = configs.tile_sample_structure(M=5, Z=1, size=(12.0, 12.0, 4.0), lam=0.4, seed=11,
                                       evanescent_fraction=0.3, incidence_deg=10.0)
    t = cw.unpack()
    plan = Plan(t)
    kz = t.W[t.begin[1]:t.end[1], 5]
    assert np.abs(kz.imag).max() * 2.0 > 10.0  # |Z| past e^10 at the slab faces
    xs, ys, zs = configs.axis(-6, 6, 16), configs.axis(-6, 6, 14), configs.axis(-1.95, 1.95, 40)
    ref = oracle.eval_grid_f64(t.W, t.z_layers, t.begin, t.end, xs, ys, zs)
    _check(Exec(plan, xs, ys, zs, form="gemm", dtype=dtype).run_host(), ref, tol, f"evanescent slab {dtype}")


@pytest.mark.parametrize("dtype", ["bf16x3", "bf16", "fp32"])
def test_wide_pupil(dtype):
    """51 x 51 orders: more than 48 kx and ky per region, so the 16-bit main
    GEMM runs with K = 128 and its in-kernel X contraction takes the K steps
    one at a time (no staged X fragments); FP32 runs the separate X
    contraction kernel."""
    cw = configs.layered_structure([1.0, 1.4 + 0.05j], [0.0], 1.0, 0.5, 51, ("x",))
    t = cw.unpack()
    plan = Plan(t)
    info = plan.region_info(0)
    assert info["separable"] and info["mx"] > 48 and info["my"] > 48
    xs, ys, zs = configs.axis(-20, 20, 37), configs.axis(-18, 18, 21), configs.axis(-5, 5, 6)
    ref = oracle.eval_grid_f64(t.W, t.z_layers, t.begin, t.end, xs, ys, zs)
    _check(Exec(plan, xs, ys, zs, form="gemm", dtype=dtype).run_host(), ref, TOL[dtype], f"wide pupil {dtype}")


@pytest.fixture(scope="module")
def stack_small():
    rng = np.random.default_rng(11)
    n = np.concatenate([[1.0], rng.uniform(1, 2, 5) + 1j * rng.uniform(0, 0.05, 5)])
    z = np.concatenate([[0.0], np.cumsum(rng.uniform(0.1, 2.0, 4))]) - 2.0
    cw = configs.layered_structure(n, z, 0.6, 0.7, 11, ("x",))
    t = cw.unpack()
    xs, ys, zs = configs.axis(-4, 4, 20), configs.axis(-4, 4, 18), configs.axis(z[0] - 1, z[-1] + 1, 40)
    ref = oracle.eval_grid_f64(t.W, t.z_layers, t.begin, t.end, xs, ys, zs)
    return Plan(t), (xs, ys, zs), ref


@pytest.mark.parametrize("form,dtype", [("direct", "fp32"), ("gemm", "fp32"), ("gemm", "bf16x3")])
def test_layer_stack(stack_small, form, dtype):
    plan, g, ref = stack_small
    _check(Exec(plan, *g, form=form, dtype=dtype).run_host(), ref, TOL[dtype] if form == "gemm" else 1e-5,
           f"layer stack {form} {dtype}")


def test_gpu_cw_evaluate_planes(c1):
    """gpu_cw_evaluate semantics: fused coordinates, region rule z >= z_l - 0.001."""
    plan, _, _ = c1
    t = plan.tables
    N, d = 32, np.float32(0.31)
    xs0, ys0, zs0 = np.float32(-4.8), np.float32(-5.1), np.float32(-4.9)
    x, y, z = 0.2, -0.4, -0.0005  # z within 1e-3 below the interface: region 1 under the GPU rule
    Exy, Exz, Eyz = gpu_cw_evaluate(plan, xs0, ys0, zs0, x, y, z, d, N, form="direct")
    from scwatter_amd.evaluator import axis_gpu
    ax, ay, az = axis_gpu(xs0, d, N), axis_gpu(ys0, d, N), axis_gpu(zs0, d, N)
    zl = t.z_layers.astype(np.float32).astype(np.float64)
    f = lambda v: [float(np.float32(v))]
    for E, grid, name in ((Exy, (ax, ay, f(z)), "xy"), (Exz, (ax, f(y), az), "xz"), (Eyz, (f(x), ay, az), "yz")):
        ref = oracle.eval_grid_f64(t.W, zl, t.begin, t.end, *grid, tol=0.001).reshape(N, N, 3)
        _check(E, ref, 1e-5, f"gpu_cw_evaluate {name}")


@pytest.mark.parametrize("kind", ["xy", "yz", "xz"])
def test_cpu_cw_evaluate_planes_match_reference_semantics(c1, kind):
    """cpu_cw_evaluate_* names run on the GPU; results agree with the FP32 reference-semantics port."""
    plan, _, _ = c1
    t = plan.tables
    N, d, a0, b0, fixed = 24, 0.41, -5.0, -4.7, -0.0005  # fixed z below the interface (CPU rule: region 0)
    E = cpu_cw_evaluate(kind, plan, a0, b0, fixed, d, N, form="direct")
    R = oracle.refcpu_plane(kind, t.W, t.z_layers, t.begin, t.end, a0, b0, fixed, d, N)
    _check(E, R, 2e-5, f"cpu_cw_evaluate_{kind} vs reference-semantics port")


def test_gemm_plane_equals_volume_slice(c1):
    plan, (xs, ys, zs), _ = c1
    V = Exec(plan, xs, ys, zs, form="gemm").run_host()
    for iz in (0, 31, 63):
        P = Exec(plan, xs, ys, [zs[iz]], form="gemm").run_host()[0]
        assert np.array_equal(P, V[iz]), f"xy plane at iz={iz} differs from the volume slice"


def test_direct_plane_matches_volume_slice(c1):
    plan, (xs, ys, zs), _ = c1
    V = Exec(plan, xs, ys, zs, form="direct").run_host()
    P = Exec(plan, xs, [ys[5]], zs, form="direct").run_host()[:, 0]
    assert oracle.rel_l2(P, V[:, 5]) < 1e-5  # both sit within FP32 rounding of the oracle


@pytest.mark.parametrize("dtype", ["bf16x3", "fp32"])
def test_half_space_with_uneven_cells(dtype):
    """A half-space region where a third of the orders lack their reflected
    wave: cells hold 1 or 2 terms, so the ELL rows are padded.  The 16-bit
    dtypes contract such narrow regions inside the slot-sum kernel, FP32 in
    the VALU coefficient kernel; both match the oracle."""
    from scwatter_amd.cw import WaveTables
    t = configs.layered_structure([1.0, 1.4 + 0.05j], [0.0], 1.0, 0.5, 9, ("x",)).unpack()
    b0, e0 = int(t.begin[0]), int(t.end[0])
    n0 = e0 - b0
    keep = np.ones(t.W.shape[0], bool)
    keep[b0 + n0 // 2: e0] = np.arange(n0 - n0 // 2) % 3 != 0  # drop every third wave of the second half
    W = t.W[keep]
    shift = np.cumsum(~keep)
    begin = np.array([b - (shift[b - 1] if b > 0 else 0) for b in t.begin], np.int32)
    end = np.array([e - shift[e - 1] for e in t.end], np.int32)
    tt = WaveTables(W, t.z_layers, begin, end)
    plan = Plan(tt)
    info = plan.region_info(0)
    assert info["separable"] and info["width"] == 2 and info["n_terms"] < 2 * info["mx"] * info["my"]
    xs, ys, zs = configs.axis(-8, 8, 20), configs.axis(-8, 8, 17), configs.axis(-4, 3, 12)
    ref = oracle.eval_grid_f64(tt.W, tt.z_layers, tt.begin, tt.end, xs, ys, zs)
    _check(Exec(plan, xs, ys, zs, form="gemm", dtype=dtype).run_host(), ref, TOL[dtype], f"uneven half-space {dtype}")


def test_empty_region_gives_zero_field():
    W = np.array([[1, 0, 0, 0.3, 0.1, 1.0]], np.complex128)
    from scwatter_amd.cw import WaveTables
    plan = Plan(WaveTables(W, np.array([0.0]), np.array([0, 1], np.int32), np.array([1, 1], np.int32)))
    for form in ("direct", "gemm"):
        E = Exec(plan, [0.0, 1.0], [0.0], [0.5, 1.5], form=form).run_host()
        assert np.all(E == 0), form


def test_output_is_overwritten_not_accumulated(c1):
    from scwatter_amd import hip
    plan, g, ref = c1
    ex = Exec(plan, *g, form="gemm")
    buf = hip.DeviceArray(ex.shape, np.complex64)
    garbage = np.full(ex.shape, 1e3 + 1e3j, np.complex64)
    buf2 = hip.DeviceArray.from_host(garbage)
    ex.run(buf2.ptr)
    ex.run(buf.ptr)
    hip.synchronize()
    assert np.array_equal(buf.to_host(), buf2.to_host())


def test_graph_replay_matches_eager(c1):
    from scwatter_amd import hip
    plan, g, _ = c1
    for form in ("direct", "gemm"):
        ex = Exec(plan, *g, form=form)
        eager = ex.run_host()
        s = hip.Stream()
        buf = hip.DeviceArray(ex.shape, np.complex64)
        ex.graph_capture(buf.ptr, s.handle)
        for _ in range(3):
            hip.DeviceArray.from_host(np.zeros(4, np.float32))  # eager work between replays
            ex.graph_launch(s.handle)
            s.synchronize()
            assert np.array_equal(buf.to_host(), eager), form
        s.destroy()


@pytest.mark.parametrize("dtype", ["bf16x3", "bf16"])
def test_graph_replay_with_both_coefficient_kernels(tile_small, dtype):
    """The tile-sample grid crosses the half-spaces (VALU term contraction) and
    the sublayers (MFMA mode contraction): eager runs and replays of the
    captured graph, with other work on the device between them, give the same
    field, which matches the oracle."""
    from scwatter_amd import hip
    plan, g, ref = tile_small
    ex = Exec(plan, *g, form="gemm", dtype=dtype)
    eager = ex.run_host()
    _check(eager, ref, TOL[dtype], f"tile sample graph {dtype}")
    s = hip.Stream()
    buf = hip.DeviceArray(ex.shape, np.complex64)
    ex.graph_capture(buf.ptr, s.handle)
    for _ in range(3):
        hip.DeviceArray.from_host(np.zeros(4, np.float32))  # eager work between replays
        ex.graph_launch(s.handle)
        s.synchronize()
        assert np.array_equal(buf.to_host(), eager)
        ex.run(buf.ptr, s.handle)  # eager run on the caller's stream
        s.synchronize()
        assert np.array_equal(buf.to_host(), eager)
    s.destroy()


def test_errors_are_reported():
    from scwatter_amd import ScwError
    from scwatter_amd.cw import WaveTables
    rng = np.random.default_rng(5)
    # polar-like pupil: every wave has its own kx and ky -> not a separable lattice
    n = 600
    W = np.column_stack([np.ones(n), np.zeros(n), np.zeros(n), rng.uniform(-1, 1, n), rng.uniform(-1, 1, n),
                         np.ones(n)])
    plan = Plan(WaveTables(W, np.zeros(0), np.array([0], np.int32), np.array([n], np.int32)))
    assert not plan.region_info(0)["separable"]
    with pytest.raises(ScwError, match="UNSUPPORTED"):
        Exec(plan, [0.0], [0.0], [0.0], form="gemm")
    assert Exec(plan, [0.0], [0.0], [0.0]).info()["form"] == "direct"
    with pytest.raises(ScwError, match="ARG"):
        Exec(plan, [], [0.0], [0.0])
    with pytest.raises(ScwError, match="ARG"):
        Exec(plan, [np.nan], [0.0], [0.0])
    bad = W.copy()
    bad[3, 0] = np.inf
    with pytest.raises(ScwError, match="ARG"):
        Plan(WaveTables(bad, np.zeros(0), np.array([0], np.int32), np.array([n], np.int32)))
    with pytest.raises(ScwError, match="ARG"):
        Plan(WaveTables(W, np.zeros(0), np.array([0], np.int32), np.array([n + 1], np.int32)))
    with pytest.raises(ScwError, match="ARG"):
        Plan(WaveTables(W, np.zeros(0), np.array([0], np.int32), np.array([n], np.int32)), device=99)


@pytest.mark.parametrize("dtype,runs", [("bf16x3", "short"), ("fp32", "short"), ("bf16x3", "long"),
                                        ("bf16", "long"), ("fp32", "long")])
def test_batch_of_uneven_instances(dtype, runs):
    """One batch run over instances of different order counts equals each
    instance against the oracle.  Short runs (9-12 z) take the 2-tile class of
    the MFMA mode contraction; long runs (33-64 z) the 8-tile class, whose
    16-bit kernel runs one workgroup per four column chunks of a run and mode
    slot (the instances have 1, 2 and 3 chunks of 64 columns, so every group
    is padded with items past the table)."""
    from scwatter_amd import BatchExec, hip
    if runs == "short":
        specs = [(3, 11, 0.6, 5.0), (5, 12, 0.9, 20.0), (7, 9, 0.45, 35.0)]
    else:
        specs = [(4, 64, 0.6, 5.0), (6, 47, 0.9, 20.0), (8, 33, 0.45, 35.0)]
    xs, ys = configs.axis(-6, 6, 20), configs.axis(-6, 6, 16)
    plans, tables, zs_list = [], [], []
    for M, nz, lam, ang in specs:
        t = configs.tile_sample_structure(M=M, Z=1, size=(12.0, 12.0, 2.0), lam=lam, seed=M,
                                          incidence_deg=ang).unpack()
        tables.append(t)
        plans.append(Plan(t))
        zs_list.append(configs.axis(-0.9, 0.9, nz))
    bx = BatchExec(plans, xs, ys, zs_list, dtype=dtype)
    assert bx.info()["form"] == "gemm"
    E = bx.run_host()
    for i, (t, zs) in enumerate(zip(tables, zs_list)):
        ref = oracle.eval_grid_f64(t.W, t.z_layers, t.begin, t.end, xs, ys, zs)
        _check(bx.instance(E, i), ref, TOL[dtype], f"batch {dtype} instance {i}")
    # a different launch shape gives the same field
    bx.set_launch(coef_items=512, main_blocks=64)
    E2 = bx.run_host()
    assert oracle.rel_l2(E2, E) < 1e-6
    s = hip.Stream()
    buf = hip.DeviceArray(bx.shape, np.complex64)
    bx.graph_capture(buf.ptr, s.handle)
    bx.graph_launch(s.handle)
    s.synchronize()
    assert np.array_equal(buf.to_host(), E2)
    s.destroy()


@pytest.mark.parametrize("dtype", ["bf16x3", "bf16", "fp16"])
def test_pupil_35_staged_x_contraction(dtype):
    """35 x 35 orders: 35 distinct kx and ky, so the 16-bit main GEMM runs
    with K = 96 and stages its X fragments in LDS over three K steps (ka =
    48), the path of every 33..48-order lattice; full grid of 14,400 points
    against the oracle."""
    t = configs.layered_structure([1.0, 1.4 + 0.05j], [0.0], 1.0, 0.5, 35, ("x",)).unpack()
    plan = Plan(t)
    for r in (0, 1):
        info = plan.region_info(r)
        assert info["separable"] and 33 <= info["mx"] <= 48 and 33 <= info["my"] <= 48
    xs, ys, zs = configs.axis(-15, 15, 40), configs.axis(-14, 16, 36), configs.axis(-6, 4, 10)
    ref = oracle.eval_grid_f64(t.W, t.z_layers, t.begin, t.end, xs, ys, zs)
    _check(Exec(plan, xs, ys, zs, form="gemm", dtype=dtype).run_host(), ref, {"bf16x3": 1e-4, "bf16": 1e-2,
                                                                             "fp16": 2e-3}[dtype],
           f"35x35 pupil gemm {dtype}")


def test_failed_relaunch_refuses_runs(tile_small):
    """A failed rebuild of the launch shape (an allocation failing part way
    through, forced with the library's allocation-failure hook) leaves the
    exec refusing every launch entry point, without launching anything; a
    later successful rebuild makes it usable again."""
    from scwatter_amd import ScwError, _native, hip
    plan, g, ref = tile_small
    ex = Exec(plan, *g, form="gemm", dtype="bf16x3")
    E0 = ex.run_host()
    L = _native.lib()
    L.scw_test_fail_alloc_after(2)  # the third allocation of the rebuild fails
    try:
        with pytest.raises(ScwError, match="NOMEM"):
            ex.set_launch(coef_items=1024, main_blocks=256)
    finally:
        L.scw_test_fail_alloc_after(-1)
    buf = hip.DeviceArray(ex.shape, np.complex64)
    s = hip.Stream()
    with pytest.raises(ScwError, match="ARG"):
        ex.run(buf.ptr)
    with pytest.raises(ScwError, match="ARG"):
        ex.run_stage(3, buf.ptr)
    with pytest.raises(ScwError, match="ARG"):
        ex.run_host()
    with pytest.raises(ScwError, match="ARG"):
        ex.graph_capture(buf.ptr, s.handle)
    ex.set_launch(coef_items=1024, main_blocks=256)
    E1 = ex.run_host()
    assert oracle.rel_l2(E1, E0) < 1e-6
    _check(E1, ref, TOL["bf16x3"], "tile sample after a failed and a good rebuild")
    s.destroy()


def test_auto_falls_back_to_direct_when_phase_overflows():
    """A 30 um slab at 0.4 um with evanescent modes: the region is separable,
    but exp(i kz (z - zref)) reaches e^90 at the slab faces, past FP32's
    range, so the GEMM form refuses this grid and AUTO takes the direct form
    instead of failing."""
    from scwatter_amd import ScwError
    cw = configs.tile_sample_structure(M=3, Z=1, size=(6.0, 6.0, 30.0), lam=0.4, seed=21,
                                       evanescent_fraction=0.5, incidence_deg=5.0)
    t = cw.unpack()
    plan = Plan(t)
    assert plan.region_info(1)["separable"]
    kz = t.W[t.begin[1]:t.end[1], 5]
    xs, ys, zs = configs.axis(-3, 3, 12), configs.axis(-3, 3, 10), configs.axis(-14.9, 14.9, 24)
    assert np.abs(kz.imag).max() * 14.9 > np.log(1e30)
    with pytest.raises(ScwError, match="UNSUPPORTED"):
        Exec(plan, xs, ys, zs, form="gemm", dtype="bf16x3")
    ex = Exec(plan, xs, ys, zs, dtype="bf16x3")
    assert ex.info()["form"] == "direct"
    ref = oracle.eval_grid_f64(t.W, t.z_layers, t.begin, t.end, xs, ys, zs)
    _check(ex.run_host(), ref, 1e-5, "thick evanescent slab, AUTO")


@pytest.mark.parametrize("orders,dtype", [(65, "bf16x3"), (65, "fp32"), (81, "bf16x3"), (81, "fp16"), (81, "fp32"),
                                          (129, "bf16x3"), (129, "bf16"), (129, "fp32")])
def test_large_pupils_chunked_main_gemm(orders, dtype):
    """81 x 81 and 129 x 129 orders: more than 64 distinct ky per region, so
    the main GEMM runs over ky chunks of at most 64 rows (two and three
    launches, each after the first adding to the output), and 129 kx take
    FP32's wide X-contraction staging (more than 128 kx).  Against the oracle
    on a grid across the interface, and the GEMM form is what AUTO picks.
    65 orders: two chunks of 33 and 32 ky, on an odd nx (rows of 6 nx floats
    that are not whole 16-byte groups, the element-wise accumulating
    stores)."""
    t = configs.layered_structure([1.0, 1.4 + 0.05j], [0.0], 1.0, 0.5, orders, ("x",)).unpack()
    plan = Plan(t)
    for r in (0, 1):
        info = plan.region_info(r)
        assert info["separable"] and info["mx"] == orders and info["my"] == orders
    xs = configs.axis(-20, 20, 31 if orders == 65 else 32)
    ys, zs = configs.axis(-18, 18, 24), configs.axis(-4, 3, 6)
    ref = oracle.eval_grid_f64(t.W, t.z_layers, t.begin, t.end, xs, ys, zs, nthreads=min(16, os.cpu_count() or 1))
    ex = Exec(plan, xs, ys, zs, dtype=dtype)
    info = ex.info()
    assert info["form"] == "gemm"
    assert info["launches"] >= (orders + 63) // 64  # one main launch per ky chunk level
    _check(ex.run_host(), ref, TOL[dtype], f"{orders}^2 pupil gemm {dtype}")



@pytest.mark.parametrize("dtype", ["bf16x3", "fp32"])
def test_tiles_change_only_the_schedule(tile_small, dtype):
    """z-run tiles (the autotuner's second family): mode-contraction runs of
    32 and 16 slices and VALU runs of 4 give the default field (the same
    sums, in other work items) and match the oracle."""
    plan, g, ref = tile_small
    ex = Exec(plan, *g, form="gemm", dtype=dtype)
    E0 = ex.run_host()
    for modal_run, coef_run in ((32, 0), (16, 4), (0, 4), (8, 1)):
        ex.set_tiles(modal_run, coef_run)
        E = ex.run_host()
        assert oracle.rel_l2(E, E0) < 1e-6, (modal_run, coef_run)
        _check(E, ref, TOL[dtype], f"tile sample tiles {modal_run}/{coef_run} {dtype}")
    with pytest.raises(Exception, match="ARG"):
        ex.set_tiles(65, 0)
