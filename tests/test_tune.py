"""The autotuner's search and cache logic on the CPU, with a stand-in batch whose run time is a
known function of its launch shape (the GPU run of the tuner is in test_gpu_tune.py)."""
import types

import pytest

from fa_amd import tune

DEFAULT = dict(lds_floats=6528, blocks_per_cu=64, use_lds=1, check_every=8, use_graph=1, tr_lds=1)


class FakeBatch:
    """set_tuning / tuning / set_timing / run as fa_amd.Batch; cost(tuning) -> seconds."""

    def __init__(self, cost, frames=()):
        self.tuning = dict(DEFAULT)
        self.cost = cost
        self.frames = list(frames)
        self._timing = False
        self.runs = 0

    def set_tuning(self, **kw):
        self.tuning = dict(self.tuning, **kw)

    def set_timing(self, on=True):
        self._timing = bool(on)


@pytest.fixture
def fake_time(monkeypatch):
    monkeypatch.setattr(tune, "time_run", lambda b, reps=3, x0=None: b.cost(b.tuning))


def test_coordinate_descent_finds_separable_minimum(fake_time):
    # blocks_per_cu best at 16, check_every best at 32, graphs better, LDS staging better
    def cost(t):
        return (1.0 + abs(t["blocks_per_cu"] - 16) / 64 + abs(t["check_every"] - 32) / 32
                + (0.2 if not t["use_graph"] else 0) + (0.3 if not t["use_lds"] else 0))
    b = FakeBatch(cost)
    best, best_t, trials = tune.tune_batch(b, reps=1, sweeps=2)
    assert best["blocks_per_cu"] == 16 and best["check_every"] == 32
    assert best["use_graph"] == 1 and best["use_lds"] == 1
    assert b.tuning == best  # the batch is left at the best shape
    assert best_t == pytest.approx(cost(best))
    assert len(trials) >= 1 + sum(len(v) - 1 for v in tune.SPACE.values())


def test_gain_threshold_keeps_current_shape(fake_time):
    # every other shape is faster by 1 %, below the 2 % gain: nothing moves
    b = FakeBatch(lambda t: 1.0 if t == DEFAULT else 0.99)
    best, _, _ = tune.tune_batch(b, reps=1, sweeps=1)
    assert best == DEFAULT
    # faster by 3 %: of equally fast candidates the first in the knob's order is taken
    b = FakeBatch(lambda t: 1.0 if t["blocks_per_cu"] == 64 else 0.97)
    best, _, _ = tune.tune_batch(b, reps=1, sweeps=1)
    assert best["blocks_per_cu"] == tune.SPACE["blocks_per_cu"][0]


def test_fastest_value_of_a_knob_wins(fake_time):
    # 8 is 3 % faster than the default, 16 faster still but by less than the gain over 8: a
    # greedy walk would stop at 8; the fastest value of the knob is taken
    times = {8: 0.97, 16: 0.955, 32: 0.99, 64: 1.0}
    b = FakeBatch(lambda t: times[t["blocks_per_cu"]])
    best, best_t, _ = tune.tune_batch(b, reps=1, sweeps=1)
    assert best["blocks_per_cu"] == 16 and best_t == pytest.approx(0.955)


def test_timing_flag_restored(fake_time):
    b = FakeBatch(lambda t: 1.0)
    b._timing = True
    tune.tune_batch(b, reps=1, sweeps=1)
    assert b._timing is True


def test_signature_and_cache_round_trip(tmp_path, fake_time):
    ref = types.SimpleNamespace(shape=(1, 64, 80), nD=9)
    ref3 = types.SimpleNamespace(shape=(4, 32, 32), nD=13)
    frames = [types.SimpleNamespace(ref=ref, dtype=1)] * 3 + [types.SimpleNamespace(ref=ref3, dtype=0)]
    ctx = types.SimpleNamespace(info=lambda: {"n_cu": 256})
    key = tune.signature(frames, ctx)
    assert key == "1x64x80/9/u16*3;4x32x32/13/f32*1@cu256"
    path = str(tmp_path / "sub" / "tune.json")
    b = FakeBatch(lambda t: 1.0 + abs(t["blocks_per_cu"] - 8) / 64, frames)
    first = tune.tuned(b, ctx, cache_path=path, reps=1, sweeps=1)
    assert first["blocks_per_cu"] == 8
    assert tune.load_cache(path)[key]["tuning"] == first
    # a second batch of the same workload takes the cached shape without timing anything
    b2 = FakeBatch(lambda t: pytest.fail("timed a cached workload"), frames)
    assert tune.tuned(b2, ctx, cache_path=path) == first and b2.tuning == first
    assert tune.load_cache(str(tmp_path / "missing.json")) == {}
