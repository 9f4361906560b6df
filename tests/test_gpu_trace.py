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
from helpers import LONG_CHAIN_CASE, MANY_NODE_CASES, PARITY_CASES, small_stack

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


LONG_CHAINS = [c for c in MANY_NODE_CASES] + [LONG_CHAIN_CASE]


def banded_solve(H, m, rhs):
    """(H) x = rhs for the block-tridiagonal H (blocks m x m): a banded LU solve, independent
    of any eigendecomposition."""
    import scipy.linalg as sl
    b = 2 * m - 1
    n = H.shape[0]
    ab = np.zeros((2 * b + 1, n))
    for d in range(-b, b + 1):
        diag = np.diagonal(H, d)
        if d >= 0:
            ab[b - d, d:] = diag
        else:
            ab[b - d, :n + d] = diag
    return sl.solve_banded((b, b), ab, rhs)


@pytest.mark.parametrize("name,shape,nD,scan,gen", LONG_CHAINS, ids=[c[0] for c in LONG_CHAINS])
def test_long_chain_steps_solve_the_subproblem(ctx, name, shape, nD, scan, gen):
    """Chains of 300 and 483 parameters (2-D and 3-D, subproblems by cyclic reduction) and 1539
    (3-D, subproblems in the batched eigen-space kernels): every traced step p with its lambda
    is the solution of
    (H + lambda I) p = -g (a banded FP64 solve of the reference's H, within 1e-9 relative), with
    H positive definite and |p| <= 1.1 radius at lambda = 0, H + lambda I positive definite and
    |p| within 0.1 radius of the radius at lambda > 0 (the subproblem's stop rule, boundary
    tolerance 0.1). Tracing does not change the results."""
    st = small_stack(shape, n_frames=3, nD=nD, seed=11, **gen)
    ref = F.Reference(ctx, st["ref_w"], nD, **scan)
    H = ref.H
    n = H.shape[0]
    ev_min = np.linalg.eigvalsh(H)[0]
    frames = [ref.register(f) for f in st["frames_w"]]
    b = F.Batch(frames, **KW)
    x_plain, _, _ = b.run()
    b.set_trace(60)
    xs, stats, _ = b.run()
    checked = 0
    for i, tg in enumerate(b.trace()):
        assert np.array_equal(xs[i], x_plain[i])
        for r in tg:
            if r[IDX["status"]] != 0:
                continue  # the frame stopped at this evaluation: no step proposed
            lam, rad = r[IDX["lambda"]], r[IDX["radius"]]
            g = r[NF + n:NF + 2 * n]
            p = r[NF + 2 * n:NF + 3 * n]
            pe = -banded_solve(H + lam * np.eye(n), ref.n_dim, g)
            assert np.linalg.norm(p - pe) <= 1e-9 * np.linalg.norm(pe), (i, r[0])
            assert r[IDX["pnorm"]] == pytest.approx(np.linalg.norm(p), rel=1e-12)
            assert r[IDX["gp"]] == pytest.approx(g @ p, rel=1e-9)
            if lam == 0:
                # the interior step, kept when it is inside the region or within the tolerance
                assert ev_min > 0 and np.linalg.norm(p) <= 1.1 * rad * (1 + 1e-12)
            else:
                assert ev_min + lam > 0
                assert abs(np.linalg.norm(p) - rad) <= 0.1 * rad * (1 + 1e-12)
            checked += 1
    assert checked >= 6, checked
    b.close()
    for f in frames:
        f.release()
    ref.release()
