"""fa_align_noalloc (the allocation-free form of fa_align) against fa_align through the C ABI:
the same result on two calls in a row, the second on a smaller reference, and a reference whose
buffers exceed the scratch array rejected with an error (checked on the host, before any launch)."""
import numpy as np
import pytest

import fa_amd as F
from helpers import PARITY_CASES, small_stack

pytestmark = pytest.mark.gpu

KW = dict(rel_delta=1e-4, max_iter=50, skip_repeat=1)


@pytest.fixture(scope="module")
def ctx():
    c = F.Context(0)
    yield c
    c.close()


def _frames(ctx, case):
    name, shape, nD, scan, gen = case
    st = small_stack(shape, nD=nD, seed=11, **gen)
    ref = F.Reference(ctx, st["ref_w"], nD, **scan)
    return ref, [ref.register(f) for f in st["frames_w"][:2]]


def test_noalloc_matches_align_larger_then_smaller(ctx):
    big = [c for c in PARITY_CASES if c[0] == "3d"][0]
    small = [c for c in PARITY_CASES if c[0] == "2d"][0]
    for case in (big, small):
        ref, frames = _frames(ctx, case)
        for fr in frames:
            x1, s1 = fr.align(**KW)
            x2, s2 = fr.align_noalloc(**KW)
            print(case[0], s1["evaluations"], s2["evaluations"], float(np.max(np.abs(x1 - x2))), flush=True)
            assert np.array_equal(x1, x2), case[0]
            assert (s1["evaluations"], s1["status"], s1["f_final"]) == (s2["evaluations"], s2["status"], s2["f_final"])
        for fr in frames:
            fr.release()
        ref.release()


def test_noalloc_long_chain(ctx):
    """A 1539-parameter chain, whose subproblems run in the batched eigen-space kernels (their
    scratch is the frame's factor scratch in the array): the same result as fa_align."""
    from helpers import LONG_CHAIN_CASE
    ref, frames = _frames(ctx, LONG_CHAIN_CASE)
    for fr in frames:
        x1, s1 = fr.align(**KW)
        x2, s2 = fr.align_noalloc(**KW)
        assert np.array_equal(x1, x2)
        assert (s1["evaluations"], s1["status"], s1["f_final"]) == (s2["evaluations"], s2["status"], s2["f_final"])
    for fr in frames:
        fr.release()
    ref.release()


def test_noalloc_reference_stop_rule_and_start(ctx):
    ref, frames = _frames(ctx, [c for c in PARITY_CASES if c[0] == "2d_interline"][0])
    x0 = np.random.default_rng(3).normal(0, 0.3, ref.nD * ref.n_dim)
    x1, s1 = frames[0].align(x0=x0)
    x2, s2 = frames[0].align_noalloc(x0=x0)
    assert np.array_equal(x1, x2) and s1["evaluations"] == s2["evaluations"]
    for fr in frames:
        fr.release()
    ref.release()


def test_noalloc_rejects_oversized_problem(ctx):
    # 1000 x 1000 cut into work items of 8 pixels: 125000 items of pass partials (8 MB)
    img = np.full((1, 1000, 1000), 10.0, np.float32)
    ref = F.Reference(ctx, img, 11, chunk_px=8)
    fr = ref.register(img)
    with pytest.raises(F.FaError, match="too large"):
        fr.align_noalloc(**KW)
    fr.release()
    ref.release()
