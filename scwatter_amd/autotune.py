"""Launch-shape and tile autotuner for the GEMM form.

Times candidate shapes on the actual exec, in full runs, keeps the fastest on
the exec, and records the choice in a JSON table keyed by the shape class of
the run, so later runs of the same class reuse it without timing
(apply_cached).  The search is coordinate-wise, one family at a time with the
best of the previous families fixed:

  1. launch shape: target coefficient work items (which sets the mode /
     term slices per region, i.e. the partial-sum slots) x target main-GEMM
     workgroups (which sets the y tiles per main workgroup);
  2. z-run tiles of the coefficient contraction: runs of up to 64, 32 or 16
     slices for the MFMA mode contraction (its 8-, 4- or 2-row-tile class) and
     of up to 8 or 4 slices for the VALU term contraction;
  3. for the 16-bit dtypes, the single-launch form (one persistent kernel per
     run) against the stage-by-stage launches.
"""
import json
import os

from . import hip

CANDIDATES = ((4096, 768), (2048, 768), (8192, 768), (16384, 768), (4096, 512), (8192, 512), (4096, 1536), (8192, 1536))
TILES = ((0, 0), (32, 0), (16, 0), (0, 4))  # (modal_run, coef_run), 0 = default (64, 8)
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


def _apply(ex, e, single=True):
    ex.set_launch(e["coef_items"], e["main_blocks"])
    if e.get("modal_run", 0) or e.get("coef_run", 0):
        ex.set_tiles(e.get("modal_run", 0), e.get("coef_run", 0))
    if single and e.get("single"):
        ex.set_persistent(True)


def apply_cached(ex, key, path=TABLE, single=True):
    """Apply the recorded shape for `key` (with single=False, the launch shape
    and tiles only, keeping the stage-by-stage form); returns its table entry
    or None."""
    e = load_table(path).get(key)
    if e:
        _apply(ex, e, single)
    return e


def _time(ex, out_ptr, stream, reps):
    e0, e1 = hip.Event(), hip.Event()
    for _ in range(2):  # the first run after a rebuild reads freshly written tables
        ex.run(out_ptr, stream.handle)
    e0.record(stream)
    for _ in range(reps):
        ex.run(out_ptr, stream.handle)
    e1.record(stream)
    e1.synchronize()
    return e0.elapsed_ms(e1) / reps


def _search(ex, out_ptr, stream, cands, setup, reps, rounds):
    """Best of `cands` (each applied by setup(c)), every candidate timed
    `rounds` times, the rounds interleaved over the candidates (so a drift of
    the clock or of the caches does not favour one), its time the best round."""
    times = {}
    for _ in range(rounds):
        for c in cands:
            setup(c)
            ms = _time(ex, out_ptr, stream, reps)
            times[c] = min(times.get(c, ms), ms)
    return min(times, key=times.get), times


def tune(ex, out_ptr, stream, reps=5, candidates=CANDIDATES, key=None, path=None, rounds=2, tiles=TILES,
         single=True):
    """Search launch shape, then tiles, then the single-launch form; keep the
    fastest on the exec.  Returns (entry, {family: {candidate: ms}}), entry =
    {coef_items, main_blocks, modal_run, coef_run, single, ms}.  With `path`,
    the entry (and every candidate's time) is merged into that JSON table
    under `key`."""
    fam = {}
    best_l, fam["launch"] = _search(ex, out_ptr, stream, tuple(candidates), lambda c: ex.set_launch(*c), reps, rounds)
    ex.set_launch(*best_l)

    def set_tiles(t):
        ex.set_tiles(*t)
    best_t, fam["tiles"] = _search(ex, out_ptr, stream, tuple(tiles), set_tiles, reps, rounds)
    ex.set_tiles(*best_t)
    best_s, ms = False, fam["tiles"][best_t]
    if single and ex.info()["dtype"] != "fp32" and ex.info()["form"] == "gemm":
        def set_single(on):
            ex.set_persistent(on)
        best_s, fam["single"] = _search(ex, out_ptr, stream, (False, True), set_single, reps, rounds)
        ex.set_persistent(best_s)
        ms = fam["single"][best_s]
    entry = {"coef_items": best_l[0], "main_blocks": best_l[1], "modal_run": best_t[0], "coef_run": best_t[1],
             "single": bool(best_s), "ms": round(ms, 4)}
    if path:
        table = load_table(path)
        table[key or shape_key(ex)] = dict(entry, candidates={
            name: {("/".join(str(x) for x in c) if isinstance(c, tuple) else str(c)): round(t, 4)
                   for c, t in times.items()} for name, times in fam.items()})
        with open(path, "w") as f:
            json.dump(table, f, indent=1, sort_keys=True)
    return entry, fam
