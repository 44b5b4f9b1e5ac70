#!/usr/bin/env python3
"""Benchmark: field pts x waves/s and rel-L2 vs the FP64 oracle, direct FP32
form vs MFMA GEMM form, on one MI355X.

Default workload (the largest configuration: 9.99e14 pts x waves per step):
  batched sweep — 64 sCWatter tile-sample slabs (one 64 x 64 x 4 um slab
  each, one internal region of 4 MF^2 eigenmode waves), lambda ~ U[0.4, 1.0]
  um, incidence ~ U[0, 40] deg, orders drawn from {11^2, 21^2, 31^2, 41^2};
  238M waves in all; each field on 256 x 256 x 64 points; every instance in
  one GEMM-form run; seeded synthetic eigen-tables.  Launch shapes come from
  the autotuner's table (scwatter_amd/tuning.json), or are tuned and recorded
  there when the table has no entry for the run's shape key.
Other workloads (--workload c1, c2, c3, c5): single interface 64^3, tile
sample 256^2 x 128, 41^2 x 2-polarisation pupil 512^2 x 256, 16-layer stack
1024^2 x 512.

A step is one evaluation of the whole grid with the tables resident in HBM.
The GEMM form is timed over K back-to-back runs, launched eagerly and as
replays of a captured hipGraph of one run (the faster counts; both are in the
line); the direct form is timed on a subset of the same workload (for the
batch: one z-slice of the first instance of each order class; its rate is
linear in the units).  Device time comes from hipEvents recorded on the
stream the kernels run on.  The CPU leg evaluates the FP64 oracle at seeded
sample points (what the rel-L2 figures are measured against) and times the
reference-semantics FP32 port (cpuEvaluator.cpp semantics) on a bounded
sample, single-threaded like the reference and on all host threads.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c4]
"""
import argparse
import re
import collections
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "field pts×waves/s + rel-L2 vs FP64 oracle: direct FP32 vs MFMA GEMM form"
UNIT = "field pts×waves/s"
PEAK = {"hbm": 8000.0, "fp32": 157.3, "bf16": 2500.0, "fp16": 2500.0}  # GB/s, TFLOP/s (MI355X, dense)
ACCURACY_BAR = 1e-4


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def time_steps(ex, out, stream, steps, warmup, single="auto"):
    """Device ms per step over `steps` back-to-back runs, launched two ways:
    eager (scw_exec_run: the stage kernels enqueued directly) and replayed from
    a captured hipGraph; with `single` "auto" or "on" (16-bit GEMM dtypes),
    the same two ways again in the single-launch form (one persistent kernel
    per run, scw_exec_set_persistent).  Returns (ms of the fastest, wall ms of
    it, which, {way: ms}) and leaves the exec in the form of the fastest way.
    On this runtime a graph replay leaves a few microseconds idle between
    consecutive replays that eager launches do not."""
    from scwatter_amd import ScwError, hip
    e0, e1 = hip.Event(), hip.Event()

    def timed(launch):
        for _ in range(warmup):
            launch()
        stream.synchronize()
        hip.synchronize()
        t0 = time.perf_counter()
        e0.record(stream)
        for _ in range(steps):
            launch()
        e1.record(stream)
        e1.synchronize()
        hip.synchronize()
        return e0.elapsed_ms(e1) / steps, (time.perf_counter() - t0) * 1e3 / steps

    res = {}
    has_single = False
    if single != "off":
        try:
            ex.set_persistent(True)
            has_single = True
        except ScwError:  # FP32 or the direct form: stage-by-stage only
            pass
    if single == "off" or single == "auto" or not has_single:
        if has_single:
            ex.set_persistent(False)
        res["eager"] = timed(lambda: ex.run(out.ptr, stream.handle))
        ex.graph_capture(out.ptr, stream.handle)
        res["graph"] = timed(lambda: ex.graph_launch(stream.handle))
    if has_single:
        ex.set_persistent(True)
        res["single_eager"] = timed(lambda: ex.run(out.ptr, stream.handle))
        ex.graph_capture(out.ptr, stream.handle)
        res["single_graph"] = timed(lambda: ex.graph_launch(stream.handle))
        ex.check(stream.handle)
    way = min(res, key=lambda k: res[k][0])
    if has_single and not way.startswith("single"):
        ex.set_persistent(False)
        ex.run(out.ptr, stream.handle)  # the output of the chosen form
        stream.synchronize()
    return res[way][0], res[way][1], way, {k: round(v[0], 4) for k, v in res.items()}


def tuned_shape(ex, out, stream, key, args):
    """Launch shape, z-run tiles and form of a GEMM-form exec from the
    autotuner's table (scwatter_amd/tuning.json) under `key`, or tuned now
    (and recorded there, and copied to --tuning-out) when the table has no
    entry or --retune is given."""
    import shutil

    from scwatter_amd import autotune
    cached = None if args.retune else autotune.apply_cached(ex, key)
    if cached:
        return {k: cached.get(k) for k in ("coef_items", "main_blocks", "modal_run", "coef_run", "single")} | {
            "source": "table", "key": key}
    entry, fam = autotune.tune(ex, out.ptr, stream, reps=4, key=key, path=autotune.TABLE)
    if args.tuning_out:
        shutil.copyfile(autotune.TABLE, args.tuning_out)
    return dict(entry, source="tuned", key=key, candidates_ms={
        name: {("/".join(str(x) for x in c) if isinstance(c, tuple) else str(c)): round(t, 4) for c, t in v.items()}
        for name, v in fam.items()})


def time_stages(ex, out, stream, reps):
    """Per-stage device time inside full runs: events between the stages of
    `reps` eager runs (coefficients, X contraction, MFMA GEMM), so every stage
    sees the caches the previous stages left, as in a real step."""
    from scwatter_amd import hip
    ev = [hip.Event() for _ in range(4)]
    tot = [0.0, 0.0, 0.0]
    for r in range(reps + 1):
        ev[0].record(stream)
        for st in (1, 2, 3):
            ex.run_stage(st, out.ptr, stream.handle)
            ev[st].record(stream)
        ev[3].synchronize()
        if r:  # the first run is a warm-up
            for st in range(3):
                tot[st] += ev[st].elapsed_ms(ev[st + 1])
    return [t / reps for t in tot]


def profiled_traffic(patterns, summary="tile_sample_c2_kernels.md"):
    """HBM bytes per step of the kernels matching `patterns` (regular
    expressions on the kernel name), from a committed rocprofv3 summary under
    profiles/ (a table of FETCH_SIZE / WRITE_SIZE per dispatch, scaled against
    a plain copy, whose last column is MB per step); None if the summary lists
    none of them."""
    path = os.path.join(ROOT, "profiles", summary)
    if not os.path.exists(path):
        return None
    total, found = 0.0, False
    for line in open(path):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) != 5:
            continue
        name = cells[0].strip("`")
        if any(re.fullmatch(p, name) for p in patterns):
            try:
                total += float(cells[4]) * 1e6
                found = True
            except ValueError:
                return None
    return total if found else None


def stage_kernels(stage, dt):
    """Kernel-name patterns of one GEMM-form stage in one dtype."""
    mode = {"fp32": 3, "bf16": 0, "bf16x3": 1, "fp16": 2}[dt]
    main = r"scw::main_f32<\d+, (true|false)>" if dt == "fp32" else r"scw::main_16<\d+, (true|false), %d>" % mode
    coef = [r"scw::coef_kernel", r"scw::modal_kernel<%d, \d>" % mode]
    if dt != "fp32":
        coef.append(r"scw::modal_wg_kernel<%d(, \d)?>" % mode)
    xc = r"scw::u_mfma" if dt == "fp32" else r"scw::csum_kernel"  # 16-bit: slot sum; X inside main_16
    return {"coef": coef, "xcontract": [xc], "main": [main]}[stage]


def write_ceiling(out, stream, reps=10):
    """GB/s of the runtime's own fill (hipMemsetAsync) over the whole output
    buffer, timed like a step: the write bandwidth the field's bytes reach
    when nothing is computed, a measured ceiling beside the 8 TB/s peak."""
    from scwatter_amd import hip
    h = hip.hip()
    e0, e1 = hip.Event(), hip.Event()
    for _ in range(2):
        hip._ok(h.hipMemsetAsync(out.ptr, 0, out.nbytes, stream.handle), "hipMemsetAsync")
    e0.record(stream)
    for _ in range(reps):
        hip._ok(h.hipMemsetAsync(out.ptr, 0, out.nbytes, stream.handle), "hipMemsetAsync")
    e1.record(stream)
    e1.synchronize()
    return out.nbytes / (e0.elapsed_ms(e1) / reps * 1e-3) / 1e9


def roofline_coef(stage, dt):
    """Executed-MFMA roofline of the coefficient stage: BF16x3 executes three
    bf16 products per algorithmic one (hi x hi, hi x lo, lo x hi), the other
    dtypes one; FP32 against the FP32 MFMA peak."""
    mult = 3.0 if dt == "bf16x3" else 1.0
    peak = PEAK["fp32"] if dt == "fp32" else PEAK["bf16"]
    a = mult * stage["flops"] / (stage["ms"] * 1e-3) / 1e12
    return {"bound": "mfma", "achieved": round(a, 2), "peak": peak, "unit": "TFLOP/s", "frac": round(a / peak, 4),
            "stage": "coef", "note": f"executed MFMA flops ({mult:g} per algorithmic flop) over the stage's "
                                     "in-pipeline time; the mode contraction and the VALU term contraction"}


def roofline(flops, nbytes, ms, peak_flops):
    t = ms * 1e-3
    t_hbm = nbytes / (PEAK["hbm"] * 1e9)
    t_mfma = flops / (peak_flops * 1e12)
    if t_hbm >= t_mfma:
        a = nbytes / t / 1e9
        return {"bound": "hbm", "achieved": round(a, 1), "peak": PEAK["hbm"], "unit": "GB/s",
                "frac": round(a / PEAK["hbm"], 4), "traffic": None}
    a = flops / t / 1e12
    return {"bound": "mfma", "achieved": round(a, 2), "peak": peak_flops, "unit": "TFLOP/s",
            "frac": round(a / peak_flops, 4), "traffic": None}


WORKLOADS = {
    "c2": ("sCWatter tile sample 21x21 orders, 20 sublayers, 15,559,803 waves; grid 256x256x128", "config2"),
    "c1": ("single interface n 1.0 / 1.4+0.05i, 21x21 focused beam, NA 0.5; grid 64^3", "config1"),
    "c3": ("single interface, 41x41 pupil x 2 polarisations (6,724 waves for z < 0, 3,362 above); grid 512x512x256 "
           "over [-60, 60]^3 um", "config3"),
    "c5": ("16-layer stack, 31x31 focused beam NA 0.7, lambda 0.6 um (33,635 waves); grid 1024x1024x512",
           "config5"),
    "c4": ("batched sweep: 64 tile-sample slabs, lambda 0.4-1.0 um, incidence 0-40 deg, orders from "
           "{11^2, 21^2, 31^2, 41^2}; 256x256x64 per instance, one batch run", "config4"),
}


def sample_device(out, shape, pidx):
    """Copy the field at the sample indices from device memory (24 bytes each)."""
    from scwatter_amd import hip
    return hip.gather_points(out.ptr, shape, pidx)


def host_info():
    """Host CPU of this box: model name and logical CPU count."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu_model": model, "nproc": os.cpu_count()}


def cpu_threads():
    """Threads for the all-core CPU legs: the host's CPUs, at most 16 (one GPU's share of a box)."""
    return max(1, min(16, os.cpu_count() or 1))


def cpu_baselines(tables, xs, ys, z, units_1, desc):
    """The reference-semantics FP32 port (cpu_cw_evaluate_xy semantics) on
    whole rows of the plane z: single-threaded on about units_1 pts x waves,
    and on all host threads on about 10 units_1, each as the median of three
    runs of a third of that.  Returns (all-thread record, single-thread
    record)."""
    import oracle
    r = oracle.region_of(float(np.float32(z)), tables.z_layers.astype(np.float32))
    w_r = int(tables.end[r] - tables.begin[r])
    out = []
    for nthr, units in ((cpu_threads(), 10 * units_1), (1, units_1)):
        # at least one row per thread (the port runs its rows in parallel),
        # each row cut to as many points as the sample leaves it
        per_row = xs.size * w_r
        units = units / 3.0
        rows = int(min(ys.size, max(nthr, round(units / max(1, per_row)))))
        npts = int(min(xs.size, max(1, round(units / (rows * w_r)))))
        xs32 = xs[:npts].astype(np.float32)
        ys32 = ys[:rows].astype(np.float32)
        zs32 = np.array([z], np.float32)
        n_units = float(npts * rows * w_r)
        times = []
        for _ in range(3):
            t0 = time.perf_counter()
            oracle.refcpu_grid(tables.W, tables.z_layers, tables.begin, tables.end, xs32, ys32, zs32, nthreads=nthr)
            times.append(time.perf_counter() - t0)
        dt = float(np.median(times))
        out.append({"value": n_units / dt, "unit": UNIT, "cores": nthr, "kind": "port",
                    "sample": f"cpu_cw_evaluate_xy semantics (FP32, {nthr} thread{'s' if nthr > 1 else ''}) on "
                              f"{rows} row(s) x {npts} points of the z={z:.3f} plane of {desc} (region {r}, "
                              f"{w_r} waves) = {n_units:.3e} units per run, median of 3 runs "
                              f"({', '.join(f'{t:.2f}' for t in times)} s)", **host_info()})
        log(f"[bench] cpu baseline ({nthr} threads): {n_units:.3e} units, median {dt:.2f} s of 3 runs -> "
            f"{n_units / dt:.3e} pts*waves/s")
    return out[0], out[1]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--direct-steps", type=int, default=1)
    ap.add_argument("--sample-points", type=int, default=4096)
    ap.add_argument("--cpu-units", type=float, default=2e8,
                    help="size of the single-thread CPU-baseline sample (pts x waves); the all-thread sample is "
                         "10x larger")
    ap.add_argument("--dtypes", default="fp32,bf16x3,bf16,fp16")
    ap.add_argument("--workload", default="c4", choices=sorted(WORKLOADS))
    ap.add_argument("--physical", action="store_true", help="tile sample (c2): eigen-tables from the host sample "
                    "solver instead of seeded synthetic ones")
    ap.add_argument("--no-drop-in", action="store_true", help="leave out the drop-in latency measurement")
    ap.add_argument("--single-launch", default="auto", choices=("auto", "on", "off"),
                    help="16-bit GEMM dtypes: also time the single-launch form and keep the faster (auto), "
                         "time only it (on), or only the stage-by-stage launches (off)")
    ap.add_argument("--tuning-out", default=None, help="after tuning, also copy the tuning table to this path")
    ap.add_argument("--retune", action="store_true", help="time the launch-shape candidates even if the "
                    "tuning table has an entry for the run")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run the HBM counters' reference kernels on 512 MiB: 5 device-to-device copies, 5 "
                         "fills and 5 min/max reductions")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: measures one GPU; --gpus {args.gpus} is not supported "
              "(the evaluation would shard by z-slab, with no exchange between GPUs).", file=sys.stderr)
        sys.exit(2)

    from scwatter_amd import Exec, Plan, autotune, configs, hip

    if args.copy_calibration:
        # reference kernels of known traffic for the HBM counters: a copy (n
        # bytes read and written), the runtime's fill (n written) and the
        # library's min/max reduction over n / 8 complex64 scalars (n read),
        # 5 each
        from scwatter_amd import post
        n = 512 << 20
        a, b = hip.DeviceArray((n,), np.uint8), hip.DeviceArray((n,), np.uint8)
        mm = hip.DeviceArray((4,), np.float32)
        stream = hip.Stream()
        for _ in range(5):
            hip._ok(hip.hip().hipMemcpy(b.ptr, a.ptr, n, hip.D2D), "hipMemcpy D2D")
        for _ in range(5):
            hip._ok(hip.hip().hipMemsetAsync(a.ptr, 0, n, stream.handle), "hipMemsetAsync")
        for _ in range(5):
            post.minmax(a.ptr, n // 8, mm.ptr, stream=stream.handle)
        stream.synchronize()
        hip.synchronize()
        print(json.dumps({"calibration_bytes": n, "copies": 5, "fills": 5, "minmax_reads": 5}), flush=True)
        return

    import oracle

    if args.workload == "c4":
        return bench_batch(args)

    desc, gen = WORKLOADS[args.workload]
    if args.physical:
        if args.workload != "c2":
            print("bench.py: --physical applies to the tile sample (--workload c2)", file=sys.stderr)
            sys.exit(2)
        desc, gen = desc + "; physical eigen-tables (host sample solver)", "config2_physical"
    t0 = time.perf_counter()
    cw, (xs, ys, zs) = getattr(configs, gen)()
    tables = cw.unpack()
    del cw
    log(f"[bench] {args.workload}: {tables.W.shape[0]} waves, {tables.layers} boundaries, grid "
        f"{xs.size}x{ys.size}x{zs.size} ({time.perf_counter()-t0:.1f} s)")
    t0 = time.perf_counter()
    plan = Plan(tables, device=0)
    log(f"[bench] plan: {plan.device_bytes()/1e9:.2f} GB on device ({time.perf_counter()-t0:.1f} s)")
    stream = hip.Stream()
    shape = (zs.size, ys.size, xs.size, 3)
    out = hip.DeviceArray(shape, np.complex64)

    # ---- seeded sample points (oracle checks) ----
    rng = np.random.default_rng(20261203)
    pidx = np.stack([rng.integers(0, n, args.sample_points) for n in (zs.size, ys.size, xs.size)], 1)
    pts = np.stack([xs[pidx[:, 2]], ys[pidx[:, 1]], zs[pidx[:, 0]]], 1)

    # ---- GEMM form, every dtype ----
    forms = {}
    wceil = write_ceiling(out, stream)
    log(f"[bench] write ceiling: hipMemsetAsync over the {out.nbytes / 1e9:.2f} GB field buffer at {wceil:.0f} GB/s")
    dtypes = [d for d in args.dtypes.split(",") if d]
    for dt in dtypes:
        t0 = time.perf_counter()
        ex = Exec(plan, xs, ys, zs, form="gemm", dtype=dt)
        tuned = tuned_shape(ex, out, stream, autotune.shape_key(ex, args.workload + ("p" if args.physical else "")),
                            args)
        info = ex.info()
        ms, wall_ms, way, by_way = time_steps(ex, out, stream, args.steps, args.warmup, args.single_launch)
        samp = sample_device(out, shape, pidx)
        stages = {}
        sms = time_stages(ex, out, stream, max(3, args.steps // 2))
        for st, name in ((1, "coef"), (2, "xcontract"), (3, "main")):
            fl, by = ex.stage_info(st)
            stages[name] = {"ms": round(sms[st - 1], 4), "flops": fl, "bytes": by}
        forms[f"gemm_{dt}"] = {"ms": ms, "wall_ms": wall_ms, "launch": way, "ms_by_launch": by_way,
                               "rate": info["units"] / (ms * 1e-3), "info": info,
                               "stages": stages, "sample": samp, "tuned": tuned}
        log(f"[bench] gemm {dt}: {ms:.3f} ms/step ({way}; {by_way}; wall {wall_ms:.3f}), "
            f"{info['units'] / (ms * 1e-3):.3e} pts*waves/s, stages " +
            ", ".join(f"{k} {v['ms']:.3f} ms" for k, v in stages.items()) + f" ({time.perf_counter()-t0:.1f} s)")
        ex.destroy()

    # ---- direct form ----
    t0 = time.perf_counter()
    exw = Exec(plan, xs[:8], ys[:8], zs[:1], form="direct")  # load the kernel
    exw.run_host()
    exw.destroy()
    ex = Exec(plan, xs, ys, zs, form="direct")
    info = ex.info()
    e0, e1 = hip.Event(), hip.Event()
    hip.synchronize()
    e0.record(stream)
    for _ in range(args.direct_steps):
        ex.run(out.ptr, stream.handle)
    e1.record(stream)
    e1.synchronize()
    dms = e0.elapsed_ms(e1) / args.direct_steps
    forms["direct_fp32"] = {"ms": dms, "rate": info["units"] / (dms * 1e-3), "info": info,
                            "sample": sample_device(out, shape, pidx)}
    log(f"[bench] direct fp32: {dms:.1f} ms/step, {info['units'] / (dms * 1e-3):.3e} pts*waves/s "
        f"({time.perf_counter()-t0:.1f} s)")
    ex.destroy()

    # ---- CPU leg: FP64 oracle on the sample points (checker) + reference-semantics FP32 baseline ----
    t0 = time.perf_counter()
    nthr = min(16, os.cpu_count() or 1)
    ref = oracle.eval_points_f64(tables.W, tables.z_layers, tables.begin, tables.end, pts, nthreads=nthr)
    log(f"[bench] oracle at {len(pts)} points ({time.perf_counter()-t0:.1f} s, {nthr} threads)")
    for k, f in forms.items():
        f["rel_l2"] = oracle.rel_l2(f["sample"], ref)
        f["max_abs"] = oracle.max_abs(f["sample"], ref)
        log(f"[bench] {k}: rel-L2 {f['rel_l2']:.3e} max-abs {f['max_abs']:.3e}")
    # CPU baseline: cpu_cw_evaluate_xy semantics on whole rows of the slice nearest z = 0.3 um
    iz = int(np.argmin(np.abs(zs - 0.3)))
    cpu_all, cpu_one = cpu_baselines(tables, xs, ys, float(zs[iz]), args.cpu_units, args.workload)

    # ---- headline: fastest GEMM dtype within the accuracy bar ----
    ok = [k for k in forms if k.startswith("gemm_") and forms[k]["rel_l2"] <= ACCURACY_BAR]
    best = max(ok, key=lambda k: forms[k]["rate"]) if ok else max(
        (k for k in forms if k.startswith("gemm_")), key=lambda k: forms[k]["rate"])
    B = forms[best]
    dt = best[len("gemm_"):]
    dom = max(B["stages"], key=lambda k: B["stages"][k]["ms"])
    S = B["stages"][dom]
    peak = PEAK["fp32"] if dt == "fp32" or dom != "main" else PEAK["bf16"]
    roof = roofline(S["flops"], S["bytes"], S["ms"], peak)
    kpat = stage_kernels(dom, dt)
    roof["kernel"] = " + ".join(kpat)
    roof["stage"] = dom
    summary = {"c2": "tile_sample_c2_kernels.md", "c5": "layer_stack_c5_kernels.md"}.get(args.workload)
    roof["traffic"] = profiled_traffic(kpat, summary) if summary else None
    roof["write_ceiling"] = round(wceil, 1)
    roof["frac_of_write_ceiling"] = round(roof["achieved"] / wceil, 4) if roof["unit"] == "GB/s" else None
    D = forms["direct_fp32"]
    roof_direct = {"bound": "valu", "achieved": round(D["info"]["flops"] / (D["ms"] * 1e-3) / 1e12, 2),
                   "peak": PEAK["fp32"], "unit": "TFLOP/s",
                   "frac": round(D["info"]["flops"] / (D["ms"] * 1e-3) / 1e12 / PEAK["fp32"], 4),
                   "note": "36 algorithmic FP32 flops + 3 transcendentals per (point, wave)"}
    line = {
        "metric": METRIC, "value": B["rate"], "unit": UNIT, "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": B["ms"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": dt, "data": "synthetic",
        "config": {"workload": desc, "units_per_step": B["info"]["units"], "form": "gemm", "gemm_dtype": dt},
        "rel_l2": B["rel_l2"], "max_abs": B["max_abs"], "accuracy_sample_points": len(pts),
        "gemm_over_direct": B["rate"] / D["rate"],
        "forms": {k: {"ms_per_step": round(f["ms"], 4), **({"launch": f["launch"], "ms_by_launch": f["ms_by_launch"]}
                                                          if "launch" in f else {}),
                      "rate": f["rate"], "rel_l2": f["rel_l2"],
                      "max_abs": f["max_abs"], **({"stages_ms": {s: v["ms"] for s, v in f["stages"].items()},
                                                   "tuned": f["tuned"]} if "stages" in f else {})}
                  for k, f in forms.items()},
        "roofline": roof, "roofline_coef": roofline_coef(B["stages"]["coef"], dt), "roofline_direct": roof_direct,
        "cpu_baseline": cpu_all, "cpu_baseline_single_thread": cpu_one,
    }
    print(json.dumps(line), flush=True)


def drop_in_latency(N=256, reps=10):
    """End-to-end latency of the reference's per-event call, gpu_cw_evaluate
    (gpuEvaluator.cu:131-165, called on every GUI event by
    sCWatterView.cpp:354-376): the three N x N planes through scw_eval_planes
    into device buffers, then copied to host arrays (the reference's three
    synchronous copies), wall time per call on the host clock.  Views of 20 um
    centred at the origin (d = 20 / (N - 1), the viewer's rule), on the tile
    sample and the 16-layer stack: the first call (execs built), the same
    call repeated (cached execs), and a drag of the xy plane by one grid step
    per call (one plane rebuilt per call), for AUTO (the GEMM form) and the
    direct form."""
    from scwatter_amd import Plan, configs, eval_planes, hip
    out = {}
    for name, gen in (("c2", configs.config2), ("c5", configs.config5)):
        t0 = time.perf_counter()
        cw, _ = gen()
        plan = Plan(cw.unpack(), device=0)
        del cw
        extent = 20.0
        d = np.float32(extent / (N - 1))
        start = np.float32(-extent / 2)
        bufs = [hip.DeviceArray((N, N, 3), np.complex64) for _ in range(3)]
        host = [np.empty((N, N, 3), np.complex64) for _ in range(3)]
        H = hip.hip()

        def call(form, zpos):
            eval_planes(plan, start, start, start, 0.1, -0.1, zpos, d, N, *(b.ptr for b in bufs), form=form)
            for hb, db in zip(host, bufs):
                hip._ok(H.hipMemcpy(hb.ctypes.data, db.ptr, hb.nbytes, hip.D2H), "hipMemcpy D2H")
        res = {}
        for form in ("auto", "direct"):
            t1 = time.perf_counter()
            call(form, 0.35)
            first = time.perf_counter() - t1
            t1 = time.perf_counter()
            for _ in range(reps):
                call(form, 0.35)
            repeat = (time.perf_counter() - t1) / reps
            t1 = time.perf_counter()
            for k in range(reps):
                call(form, 0.35 + float(d) * (k + 1))
            move = (time.perf_counter() - t1) / reps
            res[form] = {"first_ms": round(first * 1e3, 3), "repeat_ms": round(repeat * 1e3, 3),
                         "move_one_plane_ms": round(move * 1e3, 3)}
        for b in bufs:
            b.free()
        plan.destroy()
        out[name] = res
        log(f"[bench] drop-in gpu_cw_evaluate {name} N={N}: {res} ({time.perf_counter() - t0:.1f} s)")
    return {"N": N, "extent_um": 20.0, "what": "scw_eval_planes + 3 device-to-host copies, wall ms per call",
            **out}


# Oracle sample points per instance of the batch, by order count per axis:
# 21 + 15 + 12 + 16 instances of 11^2 / 21^2 / 31^2 / 41^2 orders take 4,116
# points in all, fewer on the larger instances (a point's oracle sum runs over
# all 4 MF^2 waves of its region: 58,564 ... 11,303,044).
BATCH_POINTS = {11: 100, 21: 64, 31: 48, 41: 30}


def bench_batch(args):
    """Batched sweep (config 4): every instance in one GEMM-form run, launch shape from the autotuner."""
    import oracle
    from scwatter_amd import BatchExec, Exec, Plan, autotune, configs, hip

    t0 = time.perf_counter()
    from scwatter_amd import lib
    lib()
    oracle._load()  # both loaded before the thread pool uses them
    params = configs.config4_instances()
    xs, ys, zs = configs.config4_grid()
    nzi = zs.size
    rng = np.random.default_rng(20261203)
    scale = args.sample_points / 4096.0
    samples = []  # (instance, iz, iy, ix)
    for i, p in enumerate(params):
        for _ in range(max(1, int(round(BATCH_POINTS.get(p["M"], 64) * scale)))):
            samples.append((i, int(rng.integers(nzi)), int(rng.integers(ys.size)), int(rng.integers(xs.size))))
    # the direct form's subset: the middle z-slice of the first instance of each order class
    first = {}
    for i, p in enumerate(params):
        first.setdefault(p["M"], i)
    direct_set = sorted(first.values())
    # Each instance is generated, planned and checked by the oracle at its
    # sample points (the CPU leg's checker) on a pool of 8 threads, and its
    # host tables dropped, except those of the direct subset, kept for the
    # direct form's checks and the CPU baseline.
    nthr = cpu_threads()
    pts_of = collections.defaultdict(list)
    for j, z, y, x in samples:
        pts_of[j].append([xs[x], ys[y], zs[z]])

    def build(i, t):
        plan = Plan(t, device=0)
        t2 = time.perf_counter()
        r = oracle.eval_points_f64(t.W, t.z_layers, t.begin, t.end, pts_of[i], nthreads=max(1, nthr // 8))
        return plan, r, t.W.shape[0], time.perf_counter() - t2, (t if i in direct_set else None)
    built = configs.map_instances(build, params, workers=8)
    plans = [b[0] for b in built]
    ref_parts = [b[1] for b in built]
    waves = sum(b[2] for b in built)
    t_or = sum(b[3] for b in built)
    keep = {i: b[4] for i, b in enumerate(built) if b[4] is not None}
    del built
    ref = np.concatenate(ref_parts)
    log(f"[bench] c4: {len(plans)} instances, {waves} waves, orders "
        f"{sorted(collections.Counter(p['M'] for p in params).items())}, {len(samples)} sample points "
        f"({time.perf_counter()-t0:.1f} s; oracle {t_or:.1f} thread-pool s)")
    stream = hip.Stream()
    forms = {}
    wceil = None
    for dt in [d for d in args.dtypes.split(",") if d in ("bf16x3", "fp32", "bf16", "fp16")]:
        t1 = time.perf_counter()
        bx = BatchExec(plans, xs, ys, [zs] * len(plans), dtype=dt)
        out = hip.DeviceArray(bx.shape, np.complex64)
        tuned = tuned_shape(bx, out, stream, autotune.shape_key(bx, "c4"), args)
        info = bx.info()
        ms, wall_ms, way, by_way = time_steps(bx, out, stream, args.steps, args.warmup, args.single_launch)
        sms = time_stages(bx, out, stream, 3)
        stages = {}
        for st, name in ((1, "coef"), (2, "xcontract"), (3, "main")):
            fl, by = bx.stage_info(st)
            stages[name] = {"ms": round(sms[st - 1], 4), "flops": fl, "bytes": by}
        pidx = np.array([(i * nzi + z, y, x) for i, z, y, x in samples])
        samp = sample_device(out, bx.shape, pidx)
        if wceil is None:
            wceil = write_ceiling(out, stream)
            log(f"[bench] c4 write ceiling: hipMemsetAsync over the {out.nbytes / 1e9:.2f} GB field buffer at "
                f"{wceil:.0f} GB/s")
        forms[f"gemm_{dt}"] = {"ms": ms, "wall_ms": wall_ms, "launch": way, "ms_by_launch": by_way,
                               "rate": info["units"] / (ms * 1e-3), "info": info, "stages": stages, "sample": samp,
                               "tuned": tuned}
        log(f"[bench] c4 gemm {dt}: {ms:.3f} ms/step ({way}; {by_way}), {info['units'] / (ms * 1e-3):.3e} "
            f"pts*waves/s, launch shape {tuned}, stages " +
            ", ".join(f"{k} {v['ms']:.3f} ms" for k, v in stages.items()) + f" ({time.perf_counter()-t1:.1f} s)")
        bx.destroy()
        out.free()
    for k, f in forms.items():
        f["rel_l2"] = oracle.rel_l2(f["sample"], ref)
        f["max_abs"] = oracle.max_abs(f["sample"], ref)
        log(f"[bench] c4 {k}: rel-L2 {f['rel_l2']:.3e} max-abs {f['max_abs']:.3e}")

    # ---- direct form on the subset (rate linear in units) ----
    t1 = time.perf_counter()
    iz = nzi // 2
    d_units, d_ms, d_samp, d_ref = 0.0, 0.0, [], []
    e0, e1 = hip.Event(), hip.Event()
    warm = Exec(plans[direct_set[0]], xs[:8], ys[:8], zs[:1], form="direct")
    warm.run_host()
    warm.destroy()
    for i in direct_set:
        ex = Exec(plans[i], xs, ys, zs[iz:iz + 1], form="direct")
        out = hip.DeviceArray(ex.shape, np.complex64)
        hip.synchronize()
        e0.record(stream)
        ex.run(out.ptr, stream.handle)
        e1.record(stream)
        e1.synchronize()
        d_ms += e0.elapsed_ms(e1)
        d_units += ex.info()["units"]
        t = keep[i]
        didx = np.stack([np.zeros(16, int), rng.integers(0, ys.size, 16), rng.integers(0, xs.size, 16)], 1)
        d_samp.append(sample_device(out, ex.shape, didx))
        pts = np.stack([xs[didx[:, 2]], ys[didx[:, 1]], np.full(16, zs[iz])], 1)
        d_ref.append(oracle.eval_points_f64(t.W, t.z_layers, t.begin, t.end, pts, nthreads=nthr))
        ex.destroy()
        out.free()
    d_rate = d_units / (d_ms * 1e-3)
    d_samp, d_ref = np.concatenate(d_samp), np.concatenate(d_ref)
    forms["direct_fp32"] = {"ms": d_ms, "rate": d_rate, "rel_l2": oracle.rel_l2(d_samp, d_ref),
                            "max_abs": oracle.max_abs(d_samp, d_ref), "units": d_units,
                            "subset": f"z-slice {iz} of instances {direct_set} (orders "
                                      f"{[params[i]['M'] for i in direct_set]}), 256 x 256 points each"}
    log(f"[bench] c4 direct fp32 on the subset: {d_units:.3e} units in {d_ms:.1f} ms -> {d_rate:.3e} "
        f"pts*waves/s, rel-L2 {forms['direct_fp32']['rel_l2']:.3e} ({time.perf_counter()-t1:.1f} s)")

    # ---- CPU baselines on the largest instance of the direct subset ----
    big = max(direct_set, key=lambda i: params[i]["M"])
    cpu_all, cpu_one = cpu_baselines(keep[big], xs, ys, float(zs[iz]), args.cpu_units,
                                     f"instance {big} ({params[big]['M']}^2 orders)")

    gemm = [k for k in forms if k.startswith("gemm_")]
    ok = [k for k in gemm if forms[k]["rel_l2"] <= ACCURACY_BAR]
    best = max(ok or gemm, key=lambda k: forms[k]["rate"])
    B = forms[best]
    dt = best[len("gemm_"):]
    dom = max(B["stages"], key=lambda k: B["stages"][k]["ms"])
    S = B["stages"][dom]
    peak = PEAK["fp32"] if dt == "fp32" or dom != "main" else PEAK["bf16"]
    roof = roofline(S["flops"], S["bytes"], S["ms"], peak)
    kpat = stage_kernels(dom, dt)
    roof["kernel"] = " + ".join(kpat)
    roof["stage"] = dom
    roof["traffic"] = profiled_traffic(kpat, "batched_sweep_c4_kernels.md")
    roof["write_ceiling"] = round(wceil, 1)
    roof["frac_of_write_ceiling"] = round(roof["achieved"] / wceil, 4) if roof["unit"] == "GB/s" else None
    D = forms["direct_fp32"]
    D_flops = 36.0 * D["units"]
    roof_direct = {"bound": "valu", "achieved": round(D_flops / (D["ms"] * 1e-3) / 1e12, 2), "peak": PEAK["fp32"],
                   "unit": "TFLOP/s", "frac": round(D_flops / (D["ms"] * 1e-3) / 1e12 / PEAK["fp32"], 4),
                   "note": "36 algorithmic FP32 flops + 3 transcendentals per (point, wave), on the direct subset"}
    line = {
        "metric": METRIC, "value": B["rate"], "unit": UNIT, "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": B["ms"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": dt,
        "data": "synthetic",
        "config": {"workload": WORKLOADS["c4"][0], "units_per_step": B["info"]["units"], "form": "gemm",
                   "gemm_dtype": dt, "instances": len(plans), "waves": waves},
        "rel_l2": B["rel_l2"], "max_abs": B["max_abs"], "accuracy_sample_points": len(samples),
        "gemm_over_direct": B["rate"] / D["rate"],
        "forms": {k: {"ms_per_step": round(f["ms"], 4), **({"launch": f["launch"], "ms_by_launch": f["ms_by_launch"]}
                                                          if "launch" in f else {}),
                      "rate": f["rate"], "rel_l2": f["rel_l2"], "max_abs": f["max_abs"],
                      **({"stages_ms": {s2: v["ms"] for s2, v in f["stages"].items()}, "tuned": f["tuned"]}
                         if "stages" in f else {"subset": f["subset"]})} for k, f in forms.items()},
        "roofline": roof, "roofline_coef": roofline_coef(B["stages"]["coef"], dt), "roofline_direct": roof_direct,
        "cpu_baseline": cpu_all, "cpu_baseline_single_thread": cpu_one,
        "drop_in_latency": None if args.no_drop_in else drop_in_latency(),
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
