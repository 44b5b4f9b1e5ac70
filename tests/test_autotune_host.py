"""Host logic of the autotuner (CPU): the coordinate-wise search keeps the
fastest launch shape, then the fastest z-run tiles with that shape, then the
faster of the stage-by-stage and single-launch forms; it records the choice
in the JSON table under the shape key, merges with what the table already
holds, and replays a recorded entry.  The exec and the HIP events are
stand-ins with a simulated clock: a run costs launch + tile + form terms."""
import json

import pytest

from scwatter_amd import autotune


class _Clock:
    t = 0.0


class _Event:
    def record(self, stream):
        self.t = _Clock.t

    def synchronize(self):
        pass

    def elapsed_ms(self, other):
        return other.t - self.t


class _Hip:
    Event = _Event


class _Stream:
    handle = None


class _Exec:
    shape = (8, 16, 32, 3)

    def __init__(self, launch_cost, tile_cost=None, single_cost=None, dtype="bf16x3"):
        self.lc, self.tc = launch_cost, tile_cost or {}
        self.sc = single_cost or {False: 0.0, True: 0.0}
        self.dtype = dtype
        self.launch, self.tiles, self.single, self.calls = None, (0, 0), False, []

    def info(self):
        return {"dtype": self.dtype, "form": "gemm"}

    def stage_info(self, stage):
        return (24.0 * 1234, 0.0)

    def set_launch(self, ci, mb):
        self.launch = (ci, mb)
        self.calls.append(("launch", ci, mb))

    def set_tiles(self, modal_run=0, coef_run=0):
        self.tiles = (modal_run, coef_run)
        self.calls.append(("tiles", modal_run, coef_run))

    def set_persistent(self, on=True, workgroups=0, wait_ms=0.0):
        self.single = bool(on)
        self.calls.append(("single", bool(on)))

    def run(self, out_ptr, stream):
        _Clock.t += self.lc[self.launch] + self.tc.get(self.tiles, 0.0) + self.sc[self.single]


@pytest.fixture(autouse=True)
def fake_hip(monkeypatch):
    monkeypatch.setattr(autotune, "hip", _Hip)


def test_coordinate_search_keeps_the_fastest_and_records_it(tmp_path):
    cands = ((4096, 768), (2048, 768), (8192, 512))
    tiles = ((0, 0), (32, 0), (0, 4))
    ex = _Exec({(4096, 768): 3.0, (2048, 768): 2.0, (8192, 512): 2.5}, {(32, 0): -0.5, (0, 4): 0.25},
               {False: 0.0, True: -0.1})
    table = tmp_path / "t.json"
    table.write_text(json.dumps({"other|key": {"coef_items": 1, "main_blocks": 2}}))
    entry, fam = autotune.tune(ex, 0, _Stream(), reps=4, candidates=cands, tiles=tiles, path=str(table))
    assert (entry["coef_items"], entry["main_blocks"]) == (2048, 768) and ex.launch == (2048, 768)
    assert (entry["modal_run"], entry["coef_run"]) == (32, 0) and ex.tiles == (32, 0)
    assert entry["single"] is True and ex.single
    assert entry["ms"] == pytest.approx(1.4)
    # the launch family is timed at the default tiles, the tiles at the best launch
    assert fam["launch"] == pytest.approx({(4096, 768): 3.0, (2048, 768): 2.0, (8192, 512): 2.5})
    assert fam["tiles"] == pytest.approx({(0, 0): 2.0, (32, 0): 1.5, (0, 4): 2.25})
    assert fam["single"] == pytest.approx({False: 1.5, True: 1.4})
    saved = json.loads(table.read_text())
    key = autotune.shape_key(ex)
    assert key == "bf16x3|32x16x8|terms1234|"
    s = saved[key]
    assert (s["coef_items"], s["main_blocks"], s["modal_run"], s["coef_run"], s["single"]) == (2048, 768, 32, 0, True)
    assert s["candidates"]["launch"] == {"4096/768": 3.0, "2048/768": 2.0, "8192/512": 2.5}
    assert s["candidates"]["tiles"] == {"0/0": 2.0, "32/0": 1.5, "0/4": 2.25}
    assert "other|key" in saved  # merged, not overwritten

    ex2 = _Exec({})
    assert autotune.apply_cached(ex2, key, str(table))["main_blocks"] == 768
    assert ex2.calls == [("launch", 2048, 768), ("tiles", 32, 0), ("single", True)]
    assert autotune.apply_cached(ex2, "missing", str(table)) is None and len(ex2.calls) == 3


def test_fp32_skips_the_single_launch_family(tmp_path):
    ex = _Exec({(1, 1): 1.0, (2, 2): 0.5}, dtype="fp32")
    entry, fam = autotune.tune(ex, 0, _Stream(), reps=2, candidates=((1, 1), (2, 2)), tiles=((0, 0),))
    assert entry["single"] is False and "single" not in fam
    assert not any(c[0] == "single" for c in ex.calls)


def test_default_entry_replays_only_the_launch_shape(tmp_path):
    table = tmp_path / "t.json"
    table.write_text(json.dumps({"k": {"coef_items": 8, "main_blocks": 9}}))  # older entries: launch shape only
    ex = _Exec({})
    autotune.apply_cached(ex, "k", str(table))
    assert ex.calls == [("launch", 8, 9)]


def test_missing_table_is_empty(tmp_path):
    assert autotune.load_table(str(tmp_path / "none.json")) == {}
