"""Host logic of the launch-shape autotuner (CPU): keeps the fastest
candidate on the exec, records it in the JSON table under the shape key,
merges with what the table already holds, and replays a recorded shape.
The exec and the HIP events are stand-ins with a simulated clock."""
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

    def __init__(self, cost):
        self.cost, self.launch, self.calls = cost, None, []

    def info(self):
        return {"dtype": "bf16x3"}

    def stage_info(self, stage):
        return (24.0 * 1234, 0.0)

    def set_launch(self, ci, mb):
        self.launch = (ci, mb)
        self.calls.append((ci, mb))

    def run(self, out_ptr, stream):
        _Clock.t += self.cost[self.launch]


@pytest.fixture(autouse=True)
def fake_hip(monkeypatch):
    monkeypatch.setattr(autotune, "hip", _Hip)


def test_keeps_the_fastest_and_records_it(tmp_path):
    cands = ((4096, 768), (2048, 768), (8192, 512))
    ex = _Exec({(4096, 768): 3.0, (2048, 768): 2.0, (8192, 512): 2.5})
    table = tmp_path / "t.json"
    table.write_text(json.dumps({"other|key": {"coef_items": 1, "main_blocks": 2}}))
    best, times = autotune.tune(ex, 0, _Stream(), reps=4, candidates=cands, path=str(table))
    assert best == (2048, 768) and ex.launch == (2048, 768)
    assert times == pytest.approx({(4096, 768): 3.0, (2048, 768): 2.0, (8192, 512): 2.5})
    saved = json.loads(table.read_text())
    key = autotune.shape_key(ex)
    assert key == "bf16x3|32x16x8|terms1234|"
    assert saved[key]["coef_items"] == 2048 and saved[key]["main_blocks"] == 768
    assert saved[key]["candidates"] == {"4096/768": 3.0, "2048/768": 2.0, "8192/512": 2.5}
    assert "other|key" in saved  # merged, not overwritten

    ex2 = _Exec({})
    assert autotune.apply_cached(ex2, key, str(table))["main_blocks"] == 768
    assert ex2.calls == [(2048, 768)]
    assert autotune.apply_cached(ex2, "missing", str(table)) is None and ex2.calls == [(2048, 768)]


def test_missing_table_is_empty(tmp_path):
    assert autotune.load_table(str(tmp_path / "none.json")) == {}
