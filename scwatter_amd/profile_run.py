"""Run a benchmark workload's evaluation a fixed number of times, for a profiler.

    python -m scwatter_amd.profile_run --workload c4 --dtypes bf16x3,fp32 --steps 5 --direct

builds the workload bench.py measures (same structures, grid and launch shapes:
the autotuner's table entry when there is one, in the stage-by-stage form),
runs one warm-up and --steps evaluations of the GEMM form per dtype, and with --direct the direct form
(for the batched sweep c4: one z-slice of the first instance of each order
class, as bench.py times it).  It checks nothing and times nothing itself;
it is the program to put after `rocprofv3 ... --`, where each counter pass
re-runs it.  Prints one line per step group.
"""
import argparse
import sys
import time

import numpy as np

from . import BatchExec, Exec, Plan, autotune, configs, hip, lib


def _log(msg):
    print(f"[profile_run] {msg}", flush=True)


def _shape(e):
    return {k: e[k] for k in ("coef_items", "main_blocks", "modal_run", "coef_run") if k in e} if e else "default"


def _steps(ex, out, n, what):
    t0 = time.perf_counter()
    ex.run(out.ptr)  # warm-up
    for _ in range(n):
        ex.run(out.ptr)
    hip.synchronize()
    _log(f"{what}: {n + 1} runs ({time.perf_counter() - t0:.2f} s)")


def run_batch(args):
    t0 = time.perf_counter()
    params = configs.config4_instances()
    xs, ys, zs = configs.config4_grid()
    plans = configs.map_instances(lambda i, t: Plan(t, device=0), params, workers=8)
    _log(f"c4: {len(plans)} instances planned ({time.perf_counter() - t0:.1f} s)")
    for dt in args.dtypes:
        bx = BatchExec(plans, xs, ys, [zs] * len(plans), dtype=dt)
        shape = autotune.apply_cached(bx, autotune.shape_key(bx, "c4"), single=False)
        out = hip.DeviceArray(bx.shape, np.complex64)
        _steps(bx, out, args.steps, f"c4 gemm {dt} (launch shape {_shape(shape)})")
        bx.destroy()
        out.free()
    if args.direct:
        first = {}
        for i, p in enumerate(params):
            first.setdefault(p["M"], i)
        for M, i in sorted(first.items()):
            ex = Exec(plans[i], xs, ys, zs[zs.size // 2:zs.size // 2 + 1], form="direct")
            out = hip.DeviceArray(ex.shape, np.complex64)
            ex.run(out.ptr)
            hip.synchronize()
            _log(f"c4 direct: instance {i} ({M}^2 orders), one z-slice")
            ex.destroy()
            out.free()


def run_single(args):
    gen = {"c1": configs.config1, "c2": configs.config2, "c3": configs.config3, "c5": configs.config5}[args.workload]
    t0 = time.perf_counter()
    cw, (xs, ys, zs) = gen()
    plan = Plan(cw.unpack(), device=0)
    del cw
    _log(f"{args.workload}: planned ({time.perf_counter() - t0:.1f} s)")
    out = hip.DeviceArray((zs.size, ys.size, xs.size, 3), np.complex64)
    for dt in args.dtypes:
        ex = Exec(plan, xs, ys, zs, form="gemm", dtype=dt)
        shape = autotune.apply_cached(ex, autotune.shape_key(ex, args.workload), single=False)
        _steps(ex, out, args.steps, f"{args.workload} gemm {dt} (launch shape {_shape(shape)})")
        ex.destroy()
    if args.direct:
        ex = Exec(plan, xs, ys, zs, form="direct")
        ex.run(out.ptr)
        hip.synchronize()
        _log(f"{args.workload} direct: one run")
        ex.destroy()


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--workload", default="c4", choices=["c1", "c2", "c3", "c4", "c5"])
    ap.add_argument("--dtypes", default="bf16x3")
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--direct", action="store_true")
    args = ap.parse_args(argv)
    args.dtypes = [d for d in args.dtypes.split(",") if d]
    lib()
    if hip.device_count() < 1:
        print("profile_run: no HIP device", file=sys.stderr)
        return 2
    (run_batch if args.workload == "c4" else run_single)(args)
    return 0


if __name__ == "__main__":
    sys.exit(main())
