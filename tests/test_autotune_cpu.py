"""Autotuner cache logic (tmp_amd/autotune.py) without a GPU: a stand-in
solver records what the tuner asks of it."""
import json

from tmp_amd import autotune


class FakeSolver:
    flow, ncase, nc = 2, 1, 997110

    def __init__(self):
        self.searched = 0
        self.applied = None

    def autotune(self, iters):
        self.searched += 1
        return [128, 256, 64, 128, 1]

    def set_launch_config(self, cfg):
        self.applied = list(cfg)


def test_key_rounds_shape(monkeypatch):
    monkeypatch.setattr(autotune, "device_arch", lambda: "gfx950")
    k = autotune.shape_key(FakeSolver(), 0, 1516320)
    assert k == "gfx950|launch-scheme|2|0|1|1000000|1.5"
    assert autotune._round2(23456) == 23000 and autotune._round2(7) == 7


def test_miss_searches_then_hit_applies(tmp_path, monkeypatch):
    monkeypatch.setattr(autotune, "device_arch", lambda: "gfx950")
    cache = str(tmp_path / "t.json")
    s = FakeSolver()
    blocks, hit = autotune.tune(s, 0, 1516320, cache=cache)
    assert not hit and s.searched == 1 and blocks == [128, 256, 64, 128, 1]
    assert json.load(open(cache)) == {"gfx950|launch-scheme|2|0|1|1000000|1.5": [128, 256, 64, 128, 1]}
    s2 = FakeSolver()
    blocks, hit = autotune.tune(s2, 0, 1516320, cache=cache)
    assert hit and s2.searched == 0 and s2.applied == [128, 256, 64, 128, 1]
    blocks, hit = autotune.tune(s2, 0, 1516320, cache=cache, force=True)
    assert not hit and s2.searched == 1


def test_unreadable_cache_is_a_miss(tmp_path, monkeypatch):
    monkeypatch.setattr(autotune, "device_arch", lambda: "gfx950")
    cache = tmp_path / "bad.json"
    cache.write_text("{not json")
    s = FakeSolver()
    blocks, hit = autotune.tune(s, 0, 1516320, cache=str(cache))
    assert not hit and s.searched == 1
