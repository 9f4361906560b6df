"""The trust-region trace (fa_batch_set_trace / fa_batch_get_trace) against the oracle's trace
of the same loop: on well-posed frames every evaluation of a step of at least 1e-4 px takes the
same decision (accepted or not, subproblem iterations, radius, stop) and f agrees to rounding;
tracing does not change the results. (Once the trust region has shrunk to shorter steps, as it
does near convergence, the measured decrease is as small as the rounding of f itself, and the
decisions on such steps move x by less than 1e-4 px.)"""
import numpy as np
import pytest

import oracle as O
import fa_amd as F
from helpers import PARITY_CASES, small_stack

pytestmark = pytest.mark.gpu

KW = dict(rel_delta=1e-4, max_iter=50, skip_repeat=1)
NF = len(F.TRACE_FIELDS)
IDX = {k: i for i, k in enumerate(F.TRACE_FIELDS)}


@pytest.fixture(scope="module")
def ctx():
    c = F.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("name,shape,nD,scan,gen", PARITY_CASES, ids=[c[0] for c in PARITY_CASES])
def test_trace_matches_oracle(ctx, name, shape, nD, scan, gen):
    st = small_stack(shape, nD=nD, seed=3, **gen)
    orc = O.Oracle().set_reference(st["ref_w"], nD, **scan)
    ref = F.Reference(ctx, st["ref_w"], nD, **scan)
    frames = [ref.register(f) for f in st["frames_w"]]
    b = F.Batch(frames, **KW)
    x_plain, _, _ = b.run()
    b.set_trace(200)
    xs, stats, _ = b.run()
    traces = b.trace()
    for i in range(len(frames)):
        assert np.array_equal(xs[i], x_plain[i])
        n = nD * orc.n_dim
        xo, so, to = orc.align_trace(st["frames_w"][i], cap=200, **KW)
        tg = traces[i]
        assert len(tg) == stats[i]["evaluations"] and len(to) == so["evaluations"]
        # records whose evaluated step was at least 1e-4 px long (record 0: the start point)
        m = 1
        while (m < min(len(tg), len(to)) and tg[m - 1, IDX["pnorm"]] >= 1e-4
               and to[m - 1, IDX["pnorm"]] >= 1e-4):
            m += 1
        assert m >= 3, m
        for k in ("evals", "accept", "iterations", "status", "sub_iters", "rejected", "radius"):
            assert np.array_equal(tg[:m, IDX[k]], to[:m, IDX[k]]), k
        assert np.max(np.abs(tg[:m, IDX["f_new"]] / to[:m, IDX["f_new"]] - 1)) < 1e-6
        assert np.max(np.abs(xs[i] - xo)) <= 1e-4
        # the last record's x is the returned x
        assert np.array_equal(tg[-1, NF:NF + n], xs[i])
    b.close()
    for f in frames:
        f.release()
    ref.release()
