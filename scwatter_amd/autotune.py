"""Launch-shape autotuner for the GEMM form.

Times a small set of candidate launch shapes (target coefficient work items,
target main-GEMM workgroups) on the actual exec, in full runs, keeps the
fastest on the exec, and records the choice in a JSON table keyed by the
shape class of the run, so later runs of the same class can reuse it
without timing.
"""
import json
import os

from . import hip

CANDIDATES = ((4096, 768), (2048, 768), (8192, 768), (16384, 768), (4096, 512), (8192, 512), (4096, 1536), (8192, 1536))
TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning.json")


def shape_key(ex, extra=""):
    info = ex.info()
    nz, ny, nx, _ = ex.shape
    return f"{info['dtype']}|{nx}x{ny}x{nz}|terms{int(ex.stage_info(1)[0] / 24):d}|{extra}"


def load_table(path=TABLE):
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    return {}


def apply_cached(ex, key, path=TABLE):
    """Apply a recorded launch shape for `key`; returns it or None."""
    e = load_table(path).get(key)
    if e:
        ex.set_launch(e["coef_items"], e["main_blocks"])
    return e


def tune(ex, out_ptr, stream, reps=5, candidates=CANDIDATES, key=None, path=None, rounds=2):
    """Time each candidate and keep the fastest.

    Every candidate is timed `rounds` times, the rounds interleaved over the
    candidates (so that a drift of the clock or of the caches does not favour
    one of them), each time over `reps` back-to-back runs after two warm-up
    runs (the first run after a rebuild reads freshly written tables); a
    candidate's time is its best round.  Returns (best (coef_items,
    main_blocks), {candidate: ms}).  With `path`, the result is merged into
    that JSON table under `key`.
    """
    e0, e1 = hip.Event(), hip.Event()
    times = {}
    for _ in range(rounds):
        for ci, mb in candidates:
            ex.set_launch(ci, mb)
            for _ in range(2):
                ex.run(out_ptr, stream.handle)
            e0.record(stream)
            for _ in range(reps):
                ex.run(out_ptr, stream.handle)
            e1.record(stream)
            e1.synchronize()
            ms = e0.elapsed_ms(e1) / reps
            times[(ci, mb)] = min(times.get((ci, mb), ms), ms)
    best = min(times, key=times.get)
    ex.set_launch(*best)
    if path:
        table = load_table(path)
        table[key or shape_key(ex)] = {"coef_items": best[0], "main_blocks": best[1], "ms": times[best],
                                       "candidates": {f"{c[0]}/{c[1]}": round(t, 4) for c, t in times.items()}}
        with open(path, "w") as f:
            json.dump(table, f, indent=1, sort_keys=True)
    return best, times
