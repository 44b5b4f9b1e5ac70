"""Host-side logic of bench.py (CPU): the one-GPU contract, the roofline
arithmetic and the lookup of profiled HBM traffic in the committed rocprof
summary.  No GPU call is made."""
import importlib.util
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def bench():
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_roofline_picks_the_binding_resource(bench):
    # 8e9 bytes in 2 ms: HBM time 1 ms > MFMA time of 1e9 flops -> HBM-bound at 4 TB/s
    r = bench.roofline(1e9, 8e9, 2.0, 2500.0)
    assert r["bound"] == "hbm" and r["achieved"] == pytest.approx(4000.0) and r["frac"] == pytest.approx(0.5)
    # 2.5e12 flops in 2 ms against 2.5 PFLOP/s -> MFMA-bound at half of peak
    r = bench.roofline(2.5e12, 1e6, 2.0, 2500.0)
    assert r["bound"] == "mfma" and r["frac"] == pytest.approx(0.5)


@pytest.mark.parametrize("stage", ["coef", "xcontract", "main"])
def test_profiled_traffic_found_for_every_bf16x3_stage(bench, stage):
    """Every BF16x3 stage's kernel patterns match a row of the committed
    tile-sample traffic table, so the bench line's `traffic` is not silently
    null after a kernel is renamed."""
    t = bench.profiled_traffic(bench.stage_kernels(stage, "bf16x3"))
    assert t is not None and t > 1e6


def test_stage_patterns_are_disjoint(bench):
    import re
    names = ["scw::modal_kernel<1, 1>", "scw::coef_kernel", "scw::csum_kernel", "scw::main_16<64, true, 1>",
             "scw::modal_wg_kernel<1>", "scw::modal_wg_kernel<1, 4>", "scw::modal_wg_kernel<2, 8>",
             "scw::main_f32<64, true>", "scw::u_mfma"]
    for dt in ("bf16x3", "fp32"):
        owner = {}
        for st in ("coef", "xcontract", "main"):
            for n in names:
                if any(re.fullmatch(p, n) for p in bench.stage_kernels(st, dt)):
                    assert n not in owner, (n, owner.get(n), st)
                    owner[n] = st
        assert owner.get("scw::coef_kernel") == "coef"
        if dt == "bf16x3":
            assert owner.get("scw::modal_wg_kernel<1, 4>") == "coef"


def test_coefficient_roofline_counts_executed_products(bench):
    # 1e12 algorithmic flops in 1 ms: BF16x3 executes 3e12 -> 3 PFLOP/s would pass peak; 0.5e12 -> 1.5 PF = 0.6
    r = bench.roofline_coef({"flops": 0.5e12, "ms": 1.0}, "bf16x3")
    assert r["bound"] == "mfma" and r["achieved"] == pytest.approx(1500.0) and r["frac"] == pytest.approx(0.6)
    r = bench.roofline_coef({"flops": 0.5e12, "ms": 1.0}, "fp16")
    assert r["achieved"] == pytest.approx(500.0) and r["peak"] == 2500.0
    r = bench.roofline_coef({"flops": 0.05e12, "ms": 1.0}, "fp32")
    assert r["peak"] == pytest.approx(157.3) and r["frac"] == pytest.approx(50.0 / 157.3, rel=1e-3)


def test_cpu_baseline_records(bench):
    """The cpu_baseline leg on a small sample of the single-interface
    structure: both records carry the contract's keys and name the threads
    they used, and the units their sample states are its rows x points x
    waves."""
    import re

    from scwatter_amd import configs
    cw, (xs, ys, zs) = configs.config1(grid=32)
    t = cw.unpack()
    allthr, one = bench.cpu_baselines(t, xs, ys, float(zs[-1]), 2e5, "single interface 32^3")
    assert one["cores"] == 1 and allthr["cores"] == bench.cpu_threads()
    for rec in (allthr, one):
        assert {"value", "unit", "cores", "kind", "sample"} <= set(rec)
        assert rec["kind"] == "port" and rec["unit"] == bench.UNIT and rec["value"] > 0
        m = re.search(r"(\d+) row\(s\) x (\d+) points .* (\d+) waves\) = ([0-9.e+]+) units per run", rec["sample"])
        assert m, rec["sample"]
        rows, npts, waves, units = int(m[1]), int(m[2]), int(m[3]), float(m[4])
        assert rows >= rec["cores"] and units == pytest.approx(rows * npts * waves, rel=1e-3)
