"""The single-launch form (scw_exec_set_persistent) against the stage-by-stage
runs of the same exec.

One persistent kernel takes the coefficient, slot-sum and main-GEMM units of a
run from a device queue and orders them with device counters; its arithmetic
per unit is that of the stage kernels, so its output must be bit-identical
to theirs.  The structures cover every unit kind: folded half-spaces (slot sum
+ main only, the 64^3 single-interface grid), the VALU term contraction (the
half-spaces of the tile sample, and a dense slab with the mode contraction
switched off), the mode-contraction classes of 1, 2, 4 and 8 row tiles (z
runs of 3-8, 9-16, 17-32 and 33-64 slices), K = 96 and K = 128 main GEMMs,
odd nx (unaligned rows), and grids far below the resident capacity.  The
expired-wait path is forced from the host (a slot-sum count that can never
be met) so that every kernel still ends normally."""
import numpy as np
import pytest

import oracle
from scwatter_amd import BatchExec, Exec, Plan, ScwError, configs, hip
from scwatter_amd import _native

pytestmark = pytest.mark.gpu

DTYPES16 = ["bf16x3", "bf16", "fp16"]
WAIT_MS = 500.0  # these runs take milliseconds: a wait that long means a lost count


def _log(msg):
    print(msg, flush=True)


def _same(ex, what, workgroups=0):
    """Stage-by-stage output, then the single-launch output: bit-identical."""
    ref = ex.run_host()
    ex.set_persistent(True, workgroups=workgroups, wait_ms=WAIT_MS)
    info = ex.persistent_info()
    assert info["on"] and info["units"] > 0 and 1 <= info["workgroups"] <= info["units"]
    assert ex.info()["launches"] == 1
    out = ex.run_host()
    ex.check()
    _log(f"{what}: {info['units']} units on {info['workgroups']} workgroups, "
         f"max |diff| {np.abs(out - ref).max():.3e}")
    assert np.array_equal(out.view(np.uint32), ref.view(np.uint32)), what
    return ref


@pytest.fixture(scope="module")
def c1():
    cw, g = configs.config1()
    return Plan(cw.unpack()), g


@pytest.fixture(scope="module")
def tile_small():
    cw = configs.tile_sample_structure(M=5, Z=4, size=(16.0, 16.0, 4.0), lam=1.0, seed=7)
    xs, ys, zs = configs.axis(-8, 8, 24), configs.axis(-8, 8, 20), configs.axis(-3, 3, 26)
    return Plan(cw.unpack()), (xs, ys, zs)


@pytest.mark.parametrize("dtype", DTYPES16)
def test_folded_half_spaces(c1, dtype):
    plan, g = c1
    _same(Exec(plan, *g, form="gemm", dtype=dtype), f"64^3 interface {dtype}")


@pytest.mark.parametrize("dtype", DTYPES16)
def test_tile_sample_all_unit_kinds(tile_small, dtype):
    plan, g = tile_small
    _same(Exec(plan, *g, form="gemm", dtype=dtype), f"tile sample {dtype}")


@pytest.mark.parametrize("workgroups", [1, 3, 37])
def test_small_grids(tile_small, workgroups):
    """A grid of one workgroup runs every unit itself, in table order; a few
    workgroups make most waits real."""
    plan, g = tile_small
    _same(Exec(plan, *g, form="gemm", dtype="bf16x3"), f"tile sample on {workgroups} workgroups", workgroups)


@pytest.mark.parametrize("dtype", DTYPES16)
def test_dense_slab_mode_and_term_contraction(dtype, monkeypatch):
    cw = configs.tile_sample_structure(M=7, Z=1, size=(12.0, 12.0, 2.0), lam=0.7, seed=3, incidence_deg=12.0)
    plan = Plan(cw.unpack())
    zs = np.concatenate([configs.axis(-0.95, 0.95, 75), configs.axis(1.05, 1.6, 5)])
    g = (configs.axis(-6, 6, 18), configs.axis(-6, 6, 14), zs)
    _same(Exec(plan, *g, form="gemm", dtype=dtype), f"dense slab, 64-z mode workgroups {dtype}")
    monkeypatch.setenv("SCW_MODAL_MIN_RUN", "1000000")
    _same(Exec(plan, *g, form="gemm", dtype=dtype), f"dense slab, VALU term items {dtype}")


@pytest.mark.parametrize("dtype", DTYPES16)
def test_mode_workgroup_widths(dtype, monkeypatch):
    """The 64-z mode contraction in workgroups of 4 items (256 threads) or 8
    (512 threads; the single-launch kernel runs those 4 at a time): the same
    field bit for bit either way, and the single-launch form equal to the
    stage form for both.  Three slabs of 3^2, 5^2 and 7^2 orders make groups of
    1, 2 and 3 items, so the padding to whole workgroups differs."""
    specs = [(3, 0.6, 5.0), (5, 0.9, 20.0), (7, 0.45, 35.0)]
    xs, ys = configs.axis(-6, 6, 20), configs.axis(-6, 6, 16)
    plans, zs_list = [], []
    for M, lam, ang in specs:
        t = configs.tile_sample_structure(M=M, Z=1, size=(12.0, 12.0, 2.0), lam=lam, seed=M,
                                          incidence_deg=ang).unpack()
        plans.append(Plan(t))
        zs_list.append(configs.axis(-0.9, 0.9, 64))
    outs = []
    for items in ("4", "8"):
        monkeypatch.setenv("SCW_MODAL_WG_ITEMS", items)
        outs.append(_same(BatchExec(plans, xs, ys, zs_list, dtype=dtype), f"{items}-item mode workgroups {dtype}"))
    assert np.array_equal(outs[0].view(np.uint32), outs[1].view(np.uint32))


@pytest.mark.parametrize("dtype", DTYPES16)
def test_single_slot_runs_stored_in_place(dtype, monkeypatch):
    """Mode-contraction runs with one mode slice store their sums straight
    into the slot-summed C_z the main GEMM reads, with no slot sum; with
    SCW_CS_DIRECT=0 the slot-sum kernel copies them instead.  Same field either
    way, in the stage form and the single-launch form (whose slot-sum unit of
    such a z has no blocks and only passes its count on).  Runs of 3-8, 17-32
    and 33-64 z cover the 1-, 4- and 8-tile classes."""
    specs = [(3, 0.6, 5.0, 7), (5, 0.9, 20.0, 29), (7, 0.45, 35.0, 64)]
    xs, ys = configs.axis(-6, 6, 20), configs.axis(-6, 6, 16)
    plans, zs_list = [], []
    for M, lam, ang, nz in specs:
        t = configs.tile_sample_structure(M=M, Z=1, size=(12.0, 12.0, 2.0), lam=lam, seed=M,
                                          incidence_deg=ang).unpack()
        plans.append(Plan(t))
        zs_list.append(configs.axis(-0.9, 0.9, nz))
    outs = []
    for direct in ("0", "1"):
        monkeypatch.setenv("SCW_CS_DIRECT", direct)
        bx = BatchExec(plans, xs, ys, zs_list, dtype=dtype)
        bx.set_launch(coef_items=16)  # one mode slice per region
        outs.append(_same(bx, f"single-slot runs, in place {direct}, {dtype}"))
    assert np.array_equal(outs[0], outs[1])


@pytest.mark.parametrize("nzs", [(5, 7, 8), (11, 12, 9), (20, 31, 17), (64, 47, 33)])
def test_batch_mode_classes(nzs):
    """z runs of 3-8, 9-16, 17-32 and 33-64 slices: the 1-, 2-, 4- and 8-tile
    classes of the mode contraction."""
    specs = [(3, 0.6, 5.0), (5, 0.9, 20.0), (7, 0.45, 35.0)]
    xs, ys = configs.axis(-6, 6, 20), configs.axis(-6, 6, 16)
    plans, zs_list = [], []
    for (M, lam, ang), nz in zip(specs, nzs):
        t = configs.tile_sample_structure(M=M, Z=1, size=(12.0, 12.0, 2.0), lam=lam, seed=M,
                                          incidence_deg=ang).unpack()
        plans.append(Plan(t))
        zs_list.append(configs.axis(-0.9, 0.9, nz))
    _same(BatchExec(plans, xs, ys, zs_list, dtype="bf16x3"), f"batch runs {nzs}")


@pytest.mark.parametrize("orders,dtype", [(35, "bf16x3"), (35, "fp16"), (51, "bf16x3"), (51, "bf16")])
def test_wide_pupils(orders, dtype):
    """35 orders: K = 96 with staged X fragments; 51: K = 128."""
    t = configs.layered_structure([1.0, 1.4 + 0.05j], [0.0], 1.0, 0.5, orders, ("x",)).unpack()
    g = (configs.axis(-15, 15, 40), configs.axis(-14, 16, 36), configs.axis(-6, 4, 10))
    _same(Exec(Plan(t), *g, form="gemm", dtype=dtype), f"{orders}^2 pupil {dtype}")


@pytest.mark.parametrize("nx,ny", [(45, 5), (67, 1), (3, 17)])
def test_odd_rows(tile_small, nx, ny):
    plan, _ = tile_small
    g = (configs.axis(-7, 7.5, nx), configs.axis(-6, 8, ny), configs.axis(-3, 3, 9))
    _same(Exec(plan, *g, form="gemm", dtype="bf16x3"), f"ragged {nx}x{ny}")


def test_matches_oracle(c1):
    plan, g = c1
    t = plan.tables
    ref = oracle.eval_grid_f64(t.W, t.z_layers, t.begin, t.end, *g)
    ex = Exec(plan, *g, form="gemm", dtype="bf16x3")
    ex.set_persistent(True, wait_ms=WAIT_MS)
    E = ex.run_host()
    ex.check()
    assert oracle.rel_l2(E, ref) < 1e-4


def test_graph_replays(tile_small):
    """The single launch captured in a graph: every replay (with other device
    work and eager runs between them) gives the eager field, so the counters
    are back at zero after each launch."""
    plan, g = tile_small
    ex = Exec(plan, *g, form="gemm", dtype="bf16x3")
    ex.set_persistent(True, wait_ms=WAIT_MS)
    eager = ex.run_host()
    s = hip.Stream()
    buf = hip.DeviceArray(ex.shape, np.complex64)
    ex.graph_capture(buf.ptr, s.handle)
    for _ in range(4):
        hip.DeviceArray.from_host(np.zeros(4, np.float32))  # eager work between replays
        ex.graph_launch(s.handle)
        s.synchronize()
        assert np.array_equal(buf.to_host(), eager)
        ex.run(buf.ptr, s.handle)
        s.synchronize()
        assert np.array_equal(buf.to_host(), eager)
    ex.check(s.handle)
    s.destroy()


def test_expired_wait_is_reported_and_recovered(tile_small):
    """With one count that can never be met (host-side test hook), the main
    units of z-slice 0 wait their 2 ms limit and the launch still ends;
    scw_exec_check reports SCW_ERR_TIMEOUT once.  With the count restored the
    next run is exact again: the last workgroup reset the counters."""
    plan, g = tile_small
    ex = Exec(plan, *g, form="gemm", dtype="bf16x3")
    ref = ex.run_host()
    ex.set_persistent(True, wait_ms=2.0)
    assert _native.lib().scw_test_exec_stall(ex.handle, 1) == 0
    ex.run_host()
    with pytest.raises(ScwError, match="TIMEOUT"):
        ex.check()
    ex.check()  # reported once
    assert _native.lib().scw_test_exec_stall(ex.handle, 0) == 0
    out = ex.run_host()
    ex.check()
    assert np.array_equal(out, ref)


def test_launch_shape_and_form_changes(tile_small):
    plan, g = tile_small
    ex = Exec(plan, *g, form="gemm", dtype="bf16")
    ref = ex.run_host()
    ex.set_persistent(True, wait_ms=WAIT_MS)
    ex.set_launch(coef_items=512, main_blocks=64)  # rebuilds the units too
    assert ex.persistent_info()["on"]
    E = ex.run_host()
    ex.check()
    ex.set_persistent(False)
    assert not ex.persistent_info()["on"] and ex.info()["launches"] > 1
    assert np.array_equal(ex.run_host(), E)
    assert oracle.rel_l2(E, ref) < 1e-6


def test_unsupported_forms(c1):
    plan, g = c1
    with pytest.raises(ScwError, match="UNSUPPORTED"):
        Exec(plan, *g, form="gemm", dtype="fp32").set_persistent(True)
    with pytest.raises(ScwError, match="UNSUPPORTED"):
        Exec(plan, *g, form="direct").set_persistent(True)
    ex = Exec(plan, *g, form="gemm", dtype="bf16x3")
    with pytest.raises(ScwError, match="ARG"):
        ex.set_persistent(True, workgroups=-1)
    with pytest.raises(ScwError, match="ARG"):
        ex.set_persistent(True, wait_ms=float("nan"))
    ex.check()  # stage-by-stage: nothing to report
