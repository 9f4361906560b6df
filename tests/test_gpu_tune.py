"""The autotuner on the device: every launch shape it tries gives the same results (tuning
changes where and when work runs, not what it computes), and it leaves the batch at a valid
shape of its search space."""
import numpy as np
import pytest

import fa_amd as F
from fa_amd import tune
from helpers import small_stack

pytestmark = pytest.mark.gpu

KW = dict(rel_delta=1e-4, max_iter=50, skip_repeat=1)


@pytest.fixture(scope="module")
def ctx():
    c = F.Context(0)
    yield c
    c.close()


def test_tune_batch_keeps_results(ctx):
    a = small_stack((1, 96, 80), n_frames=4, nD=9, seed=71, dtype="uint16")
    c = small_stack((6, 40, 36), n_frames=2, nD=19, seed=72, n_blobs=120, sigma=(1.5, 4.0))
    ra, rc = F.Reference(ctx, a["ref_w"], 9), F.Reference(ctx, c["ref_w"], 19)
    frames = [ra.register(f) for f in a["frames"]] + [rc.register(f) for f in c["frames_w"]]
    b = F.Batch(frames, **KW)
    x0, s0, _ = b.run()
    seen = []
    best, best_t, trials = tune.tune_batch(b, reps=1, sweeps=1, log=seen.append)
    assert best_t > 0 and len(trials) == 1 + len(seen)
    for k, v in best.items():
        if k in tune.SPACE:
            assert v in tune.SPACE[k], (k, v)
    assert b.tuning["blocks_per_cu"] == best["blocks_per_cu"]
    # every shape tried gives the same displacements and statistics
    for t, _ in trials[:: max(1, len(trials) // 4)]:
        b.set_tuning(**t)
        x1, s1, _ = b.run()
        assert all(np.array_equal(p, q) for p, q in zip(x0, x1)), t
        assert [s["evaluations"] for s in s0] == [s["evaluations"] for s in s1], t
    b.close()
    for f in frames:
        f.release()


def test_tune_chunk_picks_a_candidate(ctx):
    st = small_stack((1, 128, 128), n_frames=3, nD=11, seed=73, dtype="uint16")
    best, res = tune.tune_chunk(ctx, st["ref_w"], 11, list(st["frames"]), chunks=[2048, 4096], reps=1)
    assert best in (2048, 4096) and set(res) == {2048, 4096}
    assert all(t > 0 for t in res.values())
