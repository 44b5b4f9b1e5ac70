#!/usr/bin/env python3
"""Benchmark: field pts x waves/s and rel-L2 vs the FP64 oracle, direct FP32
form vs MFMA GEMM form, on one MI355X.

Workload (one step = one evaluation of the whole grid, tables resident in HBM):
  sCWatter 3D tile sample — 21 x 21 orders (441), 20 sublayers of 1 um between
  21 boundaries, slab 64 x 64 x 20 um at z in [-10, 10]; 2MF down + 2MF up
  eigenmodes per sublayer -> 4 MF^2 = 777,924 waves per internal region,
  15,559,803 waves in all; field on 256 x 256 x 128 points (x, y in [-32, 32],
  z in [-10, 10]); seeded synthetic eigen-tables.

The GEMM form is timed over K back-to-back runs, launched eagerly and as
replays of a captured hipGraph of one run (the faster counts; both are in the
line); the direct form (seconds per run) is timed over --direct-steps runs.  Device
time comes from hipEvents recorded on the stream the kernels run on.  The
CPU baseline is the reference-semantics FP32 port (single thread, like
cpuEvaluator.cpp) on a bounded sample of the same workload; the same CPU
leg evaluates the FP64 oracle on seeded sample points, which is what the
rel-L2 figures are measured against.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--direct-steps D]
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


def time_steps(ex, out, stream, steps, warmup):
    """Device ms per step over `steps` back-to-back runs, launched two ways:
    eager (scw_exec_run: the stage kernels enqueued directly) and replayed from
    a captured hipGraph.  Returns (ms of the faster, wall ms of it, which,
    {way: ms}).  On this runtime a graph replay leaves a few microseconds
    idle between consecutive replays that eager launches do not."""
    from scwatter_amd import hip
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

    res = {"eager": timed(lambda: ex.run(out.ptr, stream.handle))}
    ex.graph_capture(out.ptr, stream.handle)
    res["graph"] = timed(lambda: ex.graph_launch(stream.handle))
    way = min(res, key=lambda k: res[k][0])
    return res[way][0], res[way][1], way, {k: round(v[0], 4) for k, v in res.items()}


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


def profiled_traffic(patterns):
    """HBM bytes per step of the kernels matching `patterns` (regular
    expressions on the kernel name), from the committed rocprofv3 summary of
    the tile-sample workload (FETCH_SIZE / WRITE_SIZE per dispatch, scaled
    against a plain copy); None if the summary lists none of them."""
    path = os.path.join(ROOT, "profiles", "tile_sample_c2_kernels.md")
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
        coef.append(r"scw::modal_wg_kernel<%d>" % mode)
    xc = r"scw::u_mfma" if dt == "fp32" else r"scw::csum_kernel"  # 16-bit: slot sum; X inside main_16
    return {"coef": coef, "xcontract": [xc], "main": [main]}[stage]


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
    res = np.empty((len(pidx), 3), np.complex64)
    nz, ny, nx, _ = shape
    for i, (iz, iy, ix) in enumerate(pidx):
        off = ((int(iz) * ny + int(iy)) * nx + int(ix)) * 24
        hip._ok(hip.hip().hipMemcpy(res[i].ctypes.data, out.ptr + off, 24, hip.D2H), "hipMemcpy D2H")
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--direct-steps", type=int, default=1)
    ap.add_argument("--sample-points", type=int, default=192)
    ap.add_argument("--cpu-units", type=float, default=2e8, help="size of the CPU-baseline sample (pts x waves)")
    ap.add_argument("--dtypes", default="fp32,bf16x3,bf16,fp16")
    ap.add_argument("--workload", default="c2", choices=sorted(WORKLOADS))
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run 5 device-to-device copies of 512 MiB (reference kernel for HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: measures one GPU; --gpus {args.gpus} is not supported "
              "(the evaluation would shard by z-slab, with no exchange between GPUs).", file=sys.stderr)
        sys.exit(2)

    from scwatter_amd import Exec, Plan, configs, hip

    if args.copy_calibration:
        n = 512 << 20
        a, b = hip.DeviceArray((n,), np.uint8), hip.DeviceArray((n,), np.uint8)
        for _ in range(5):
            hip._ok(hip.hip().hipMemcpy(b.ptr, a.ptr, n, hip.D2D), "hipMemcpy D2D")
        hip.synchronize()
        print(json.dumps({"copy_calibration_bytes": n, "copies": 5}), flush=True)
        return

    import oracle

    if args.workload == "c4":
        return bench_batch(args)

    desc, gen = WORKLOADS[args.workload]
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
    dtypes = [d for d in args.dtypes.split(",") if d]
    for dt in dtypes:
        t0 = time.perf_counter()
        ex = Exec(plan, xs, ys, zs, form="gemm", dtype=dt)
        info = ex.info()
        ms, wall_ms, way, by_way = time_steps(ex, out, stream, args.steps, args.warmup)
        samp = sample_device(out, shape, pidx)
        stages = {}
        sms = time_stages(ex, out, stream, max(3, args.steps // 2))
        for st, name in ((1, "coef"), (2, "xcontract"), (3, "main")):
            fl, by = ex.stage_info(st)
            stages[name] = {"ms": round(sms[st - 1], 4), "flops": fl, "bytes": by}
        forms[f"gemm_{dt}"] = {"ms": ms, "wall_ms": wall_ms, "launch": way, "ms_by_launch": by_way,
                               "rate": info["units"] / (ms * 1e-3), "info": info,
                               "stages": stages, "sample": samp}
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
    # CPU baseline: cpu_cw_evaluate_xy semantics on whole rows of the slice nearest z = 0.3 um,
    # as many rows as make about --cpu-units units.
    iz = int(np.argmin(np.abs(zs - 0.3)))
    r = oracle.region_of(float(np.float32(zs[iz])), tables.z_layers.astype(np.float32))
    w_r = int(tables.end[r] - tables.begin[r])
    rows = int(max(1, min(ys.size, round(args.cpu_units / max(1, xs.size * w_r)))))
    xs32, ys32, zs32 = xs.astype(np.float32), ys[:rows].astype(np.float32), zs[iz:iz + 1].astype(np.float32)
    cpu_units = float(xs.size * rows * w_r)
    t0 = time.perf_counter()
    oracle.refcpu_grid(tables.W, tables.z_layers, tables.begin, tables.end, xs32, ys32, zs32, nthreads=1)
    cpu_s = time.perf_counter() - t0
    cpu_rate = cpu_units / cpu_s
    log(f"[bench] cpu baseline: {cpu_units:.3e} units in {cpu_s:.1f} s -> {cpu_rate:.3e} pts*waves/s")

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
    roof["traffic"] = profiled_traffic(kpat) if args.workload == "c2" else None
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
                      "max_abs": f["max_abs"], **({"stages_ms": {s: v["ms"] for s, v in f["stages"].items()}}
                                                  if "stages" in f else {})} for k, f in forms.items()},
        "roofline": roof, "roofline_direct": roof_direct,
        "cpu_baseline": {"value": cpu_rate, "unit": UNIT, "cores": 1, "kind": "port",
                         "sample": f"cpu_cw_evaluate_xy semantics (FP32, 1 thread) on {rows} row(s) x "
                                   f"{xs.size} points of the z={zs[iz]:.3f} slice (region {r}, {w_r} waves) = "
                                   f"{cpu_units:.3e} units"},
    }
    print(json.dumps(line), flush=True)


def bench_batch(args):
    """Batched sweep (config 4): every instance in one GEMM-form run, launch shape from the autotuner."""
    import oracle
    from scwatter_amd import BatchExec, Plan, autotune, configs, hip

    t0 = time.perf_counter()
    params = configs.config4_instances()
    xs, ys, zs = configs.config4_grid()
    nzi = zs.size
    rng = np.random.default_rng(20261203)
    samples = [(i, int(rng.integers(nzi)), int(rng.integers(ys.size)), int(rng.integers(xs.size)))
               for i in range(len(params)) for _ in range(3)]
    # Each instance is planned, checked by the oracle at its 3 sample points
    # (the CPU leg's checker), and its host tables dropped, except the
    # largest instance's, kept for the CPU baseline.
    plans, ref_parts, waves, big = [], [], 0, None
    nthr = min(16, os.cpu_count() or 1)
    t_or = 0.0
    for i, p in enumerate(params):
        t = configs.config4_structure(p).unpack()
        plans.append(Plan(t, device=0))
        waves += t.W.shape[0]
        t2 = time.perf_counter()
        pts = [[xs[x], ys[y], zs[z]] for j, z, y, x in samples if j == i]
        ref_parts.append(oracle.eval_points_f64(t.W, t.z_layers, t.begin, t.end, pts, nthreads=nthr))
        t_or += time.perf_counter() - t2
        if big is None or t.W.shape[0] > big.W.shape[0]:
            big = t
        del t
        if i % 8 == 7:
            log(f"[bench] c4: {i + 1} instances planned ({time.perf_counter()-t0:.1f} s)")
    ref = np.concatenate(ref_parts)
    log(f"[bench] c4: {len(plans)} instances, {waves} waves, orders "
        f"{sorted(collections.Counter(p['M'] for p in params).items())} ({time.perf_counter()-t0:.1f} s, "
        f"oracle {t_or:.1f} s)")
    stream = hip.Stream()
    forms = {}
    for dt in [d for d in args.dtypes.split(",") if d in ("bf16x3", "fp32", "bf16", "fp16")]:
        t1 = time.perf_counter()
        bx = BatchExec(plans, xs, ys, [zs] * len(plans), dtype=dt)
        out = hip.DeviceArray(bx.shape, np.complex64)
        best, cand = autotune.tune(bx, out.ptr, stream, reps=3)
        info = bx.info()
        ms, wall_ms, way, by_way = time_steps(bx, out, stream, args.steps, args.warmup)
        sms = time_stages(bx, out, stream, 3)
        stages = {}
        for st, name in ((1, "coef"), (2, "xcontract"), (3, "main")):
            fl, by = bx.stage_info(st)
            stages[name] = {"ms": round(sms[st - 1], 4), "flops": fl, "bytes": by}
        pidx = np.array([(i * nzi + z, y, x) for i, z, y, x in samples])
        samp = sample_device(out, bx.shape, pidx)
        forms[f"gemm_{dt}"] = {"ms": ms, "wall_ms": wall_ms, "launch": way, "ms_by_launch": by_way,
                               "rate": info["units"] / (ms * 1e-3), "info": info,
                               "stages": stages, "sample": samp, "tuned": {"coef_items": best[0], "main_blocks": best[1],
                                                                           "candidates_ms": {f"{a}/{b}": round(v, 4)
                                                                                             for (a, b), v in cand.items()}}}
        log(f"[bench] c4 gemm {dt}: {ms:.3f} ms/step, {info['units'] / (ms * 1e-3):.3e} pts*waves/s, tuned {best}, "
            "stages " + ", ".join(f"{k} {v['ms']:.3f} ms" for k, v in stages.items()) +
            f" ({time.perf_counter()-t1:.1f} s)")
        bx.destroy()
        out.free()
    for k, f in forms.items():
        f["rel_l2"] = oracle.rel_l2(f["sample"], ref)
        f["max_abs"] = oracle.max_abs(f["sample"], ref)
        log(f"[bench] c4 {k}: rel-L2 {f['rel_l2']:.3e} max-abs {f['max_abs']:.3e}")
    # CPU baseline: reference-semantics FP32 port on a bounded sample of the largest instance
    tb = big
    w_r = int(tb.end[1] - tb.begin[1])
    npts = int(max(1, min(xs.size, round(args.cpu_units / w_r))))
    t1 = time.perf_counter()
    oracle.refcpu_grid(tb.W, tb.z_layers, tb.begin, tb.end, xs[:npts].astype(np.float32),
                       ys[:1].astype(np.float32), zs[nzi // 2:nzi // 2 + 1].astype(np.float32), nthreads=1)
    cpu_s = time.perf_counter() - t1
    cpu_units = float(npts * w_r)
    ok = [k for k in forms if forms[k]["rel_l2"] <= ACCURACY_BAR]
    best = max(ok or list(forms), key=lambda k: forms[k]["rate"])
    B = forms[best]
    dt = best[len("gemm_"):]
    dom = max(B["stages"], key=lambda k: B["stages"][k]["ms"])
    S = B["stages"][dom]
    roof = roofline(S["flops"], S["bytes"], S["ms"], PEAK["fp32"] if dt == "fp32" else PEAK["bf16"])
    roof["kernel"] = " + ".join(stage_kernels(dom, dt))
    roof["stage"] = dom
    line = {
        "metric": METRIC, "value": B["rate"], "unit": UNIT, "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": B["ms"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": dt,
        "data": "synthetic",
        "config": {"workload": WORKLOADS["c4"][0], "units_per_step": B["info"]["units"], "form": "gemm",
                   "gemm_dtype": dt, "instances": len(plans), "waves": waves},
        "rel_l2": B["rel_l2"], "max_abs": B["max_abs"], "accuracy_sample_points": len(samples),
        "forms": {k: {"ms_per_step": round(f["ms"], 4), **({"launch": f["launch"], "ms_by_launch": f["ms_by_launch"]}
                                                          if "launch" in f else {}),
                      "rate": f["rate"], "rel_l2": f["rel_l2"],
                      "max_abs": f["max_abs"], "stages_ms": {s: v["ms"] for s, v in f["stages"].items()},
                      "tuned": f["tuned"]} for k, f in forms.items()},
        "roofline": roof,
        "cpu_baseline": {"value": cpu_units / cpu_s, "unit": UNIT, "cores": 1, "kind": "port",
                         "sample": f"cpu_cw_evaluate_xy semantics (FP32, 1 thread) on {npts} points of one row of the "
                                   f"largest instance ({w_r} waves) = {cpu_units:.3e} units"},
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
