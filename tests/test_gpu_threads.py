"""Per-frame calls from several host threads (the reference's MEX pool runs addFrame on
hardware_concurrency threads, FrameAlignerMex.cpp:24-25,120): calls on different frames run at
the same time, each on its frame's own batch and stream, and give the results of the same
calls made one after the other."""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import fa_amd as F
from helpers import small_stack

pytestmark = pytest.mark.gpu

KW = dict(rel_delta=1e-4, max_iter=50, skip_repeat=1)


@pytest.fixture(scope="module")
def ctx():
    c = F.Context(0)
    yield c
    c.close()


def test_concurrent_align_and_evaluations(ctx):
    a = small_stack((1, 96, 80), n_frames=6, nD=9, seed=61)
    b = small_stack((4, 40, 36), n_frames=3, nD=13, seed=62, n_blobs=120, sigma=(1.5, 4.0))
    ra, rb = F.Reference(ctx, a["ref_w"], 9), F.Reference(ctx, b["ref_w"], 13)
    frames = [ra.register(f) for f in a["frames_w"]] + [rb.register(f) for f in b["frames_w"]]
    alone = [f.align(**KW) for f in frames]
    rng = np.random.default_rng(3)
    Ds = [rng.normal(0, 0.5, (f.ref.nD, 3)) for f in frames]
    ev_alone = [f.jacobian(D) for f, D in zip(frames, Ds)]
    with ThreadPoolExecutor(len(frames)) as ex:
        for _ in range(2):
            got = list(ex.map(lambda f: f.align(**KW), frames))
            ev = list(ex.map(lambda fd: fd[0].jacobian(fd[1]), zip(frames, Ds)))
            for i, ((x1, s1), (x2, s2)) in enumerate(zip(alone, got)):
                assert np.array_equal(x1, x2), i
                assert (s1["evaluations"], s1["status"], s1["f_final"]) == (s2["evaluations"], s2["status"], s2["f_final"]), i
            for i, ((e1, j1), (e2, j2)) in enumerate(zip(ev_alone, ev)):
                assert e1 == e2 and np.array_equal(j1, j2), i
    # the same frame from two threads: the calls take turns
    with ThreadPoolExecutor(2) as ex:
        got = list(ex.map(lambda _: frames[0].align(**KW), range(2)))
    assert all(np.array_equal(x, alone[0][0]) for x, _ in got)
    for f in frames:
        f.release()
