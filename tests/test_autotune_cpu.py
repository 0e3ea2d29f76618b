"""The schedule autotuner's search and cache, on a stand-in object whose step time is a known
function of the schedule (no GPU)."""
import json

from bg_flood_amd import autotune


class FakeStep:
    """set_schedule/schedule like a Domain; run(n) 'takes' a time that the test fixes per schedule."""

    def __init__(self, cost):
        self.cost = cost
        self.s = (16, 12, 0, 0)
        self.clock = 0.0

    def set_schedule(self, tx=0, ty=0, wp=0, wc=0):
        self.s = (tx or self.s[0], ty or self.s[1], wp, wc)

    def schedule(self):
        return dict(zip(("tile_x", "tile_y", "wg_pred", "wg_corr"), self.s))

    def run(self, n):
        self.clock += n * self.cost(*self.s)


def test_tune_finds_the_coordinate_minimum(monkeypatch):
    def cost(tx, ty, wp, wc):
        c = 1.0 + (0.0 if (tx, ty) == (32, 8) else 0.5)
        c += {0: 0.3, 1: 0.4, 2: 0.0, 3: 0.2}.get(wp, 0.3)
        c += {0: 0.2, 1: 0.0}.get(wc, 0.2)
        return c

    f = FakeStep(cost)
    monkeypatch.setattr(autotune.time, "perf_counter", lambda: f.clock)
    best, trials = autotune.tune(f, f.run, steps=8, repeats=2, max_wg=4)
    assert best == dict(tile_x=32, tile_y=8, wg_pred=2, wg_corr=1)
    assert f.s == (32, 8, 2, 1)  # left applied
    assert len(trials) == len(autotune.TILES) + 4 + 4
    for t in trials:
        assert abs(t["s_per_step"] - cost(t["tile_x"], t["tile_y"], t["wg_pred"], t["wg_corr"])) < 1e-12


def test_tune_keeps_the_default_on_ties(monkeypatch):
    f = FakeStep(lambda *s: 1.0)
    monkeypatch.setattr(autotune.time, "perf_counter", lambda: f.clock)
    best, _ = autotune.tune(f, f.run, steps=4, repeats=1, max_wg=2)
    assert best == dict(tile_x=16, tile_y=12, wg_pred=0, wg_corr=0)


def test_cache_round_trip(tmp_path):
    path = str(tmp_path / "a" / "s.json")
    k = autotune.cache_key("batch", "fp32_comp", 1000, "64")
    assert k == "batch|fp32_comp|1024|64"
    assert autotune.cache_key("domain", "fp32", 1024) == "domain|fp32|1024"
    autotune.store_cache(path, k, dict(tile_x=8, tile_y=8, wg_pred=1, wg_corr=2))
    autotune.store_cache(path, "x", dict(tile_x=1, tile_y=1, wg_pred=0, wg_corr=0))
    data = json.load(open(path))
    assert data[k] == dict(tile_x=8, tile_y=8, wg_pred=1, wg_corr=2) and "x" in data
    assert autotune.load_cache(str(tmp_path / "none.json")) == {}
