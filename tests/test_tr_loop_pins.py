"""Pins of the oracle's trust-region control flow (dlib find_min_trust_region with
objective_delta_stop_strategy, as FrameWarpAligner.cpp:205-209 calls it) on closed-form models,
through oracle.tr_run (orc_tr_run: the same loop orc_align runs over the warper).

In one dimension everything the loop decides can be derived by hand. With f = a (x - c)^2 / 2,
a model gradient g = s a (x - c) (s = 1: the true gradient; s != 1 stands in for the warper's
Gauss-Newton Jtr, which is not the exact derivative) and B = b:
  * an interior step (|g| / b <= radius) is p = -g / b, found in one subproblem iteration, and
    rho = (2 - s a / b) / s;
  * a boundary step (|g| / b > radius) has |p| = radius, found in two iterations (Newton on
    1/|p(lambda)| - 1/radius is exact after one update in 1-D), and with d = |x - c|
    rho = (a d - a R / 2) / (s a d - b R / 2).
Choosing s puts rho on either side of each threshold, so a wrong threshold (0.25, 0.75), factor
(1/4, 2), cap (1000), expansion rule (only for steps the radius held), stop rule or stale-step
handling fails one of these tests. What they do not pin is the lambda safeguarding inside
dlib's solve_trust_region_subproblem in several dimensions (dlib is not available here).
"""
import numpy as np
import pytest

import oracle as O

IDX = {k: i for i, k in enumerate(O.TRACE_FIELDS)}
EPS = np.finfo(float).eps


def run1d(a, b, c, s, x0, cap=200, **kw):
    f = lambda x: 0.5 * a * (x[0] - c) ** 2
    g = lambda x: np.array([s * a * (x[0] - c)])
    return O.tr_run(np.array([[[b]]]), np.zeros((0, 1, 1)), f, g, [x0], cap=cap, **kw)


def col(tr, k):
    return tr[:, IDX[k]]


def s_interior(rho, a=1.0, b=1.0):
    """s giving an interior step of ratio rho: (2 - s a/b)/s = rho."""
    return 2.0 / (rho + a / b)


def s_boundary(rho, d, R, a=1.0, b=1.0):
    """s giving a boundary step of ratio rho at distance d with radius R."""
    return ((a * d - 0.5 * a * R) / rho + 0.5 * b * R) / (a * d)


def test_subproblem_iterations_and_rho_closed_forms():
    # interior: p = -g/b in one iteration; boundary: |p| = R in two
    for s, x0, interior in ((s_interior(1.3), 10.0, True), (s_boundary(1.3, 1000.0, 40.0), 1000.0, False)):
        x, st, tr = run1d(1.0, 1.0, 0.0, s, x0, max_evals=2)
        p = tr[0, O.N_TRACE + 2]
        if interior:
            assert tr[0, IDX["sub_iters"]] == 1 and tr[0, IDX["lambda"]] == 0
            assert p == pytest.approx(-s * x0)
        else:
            assert tr[0, IDX["sub_iters"]] == 2
            assert p == pytest.approx(-40.0, rel=1e-12)
        assert tr[1, IDX["rho"]] == pytest.approx(1.3, rel=1e-12)
        # pred = -(g p + b p^2 / 2), meas = f(x) - f(x + p)
        g0 = s * x0
        assert tr[1, IDX["pred"]] == pytest.approx(-(g0 * p + 0.5 * p * p), rel=1e-12)
        assert tr[1, IDX["meas"]] == pytest.approx(0.5 * x0 ** 2 - 0.5 * (x0 + p) ** 2, rel=1e-12)


@pytest.mark.parametrize("rho,expanded", [(0.76, True), (0.74, False)])
def test_expansion_threshold_on_boundary_steps(rho, expanded):
    s = s_boundary(rho, 1000.0, 40.0)
    _, _, tr = run1d(1.0, 1.0, 0.0, s, 1000.0, max_evals=2)
    assert tr[0, IDX["sub_iters"]] == 2
    assert tr[1, IDX["rho"]] == pytest.approx(rho, rel=1e-10)
    assert tr[1, IDX["radius"]] == (80.0 if expanded else 40.0)
    assert tr[1, IDX["accept"]] == 1


def test_no_expansion_after_interior_step():
    # rho = 0.9 > 0.75, but the step was interior (one subproblem iteration): radius stays
    _, _, tr = run1d(1.0, 1.0, 0.0, s_interior(0.9), 10.0, max_evals=2)
    assert tr[0, IDX["sub_iters"]] == 1
    assert tr[1, IDX["rho"]] == pytest.approx(0.9, rel=1e-12)
    assert tr[1, IDX["radius"]] == 40.0


@pytest.mark.parametrize("rho,radius", [(0.26, 40.0), (0.24, 10.0), (0.01, 10.0)])
def test_shrink_threshold_and_factor(rho, radius):
    # 0 < rho < 0.25: radius / 4 and the step is still taken
    x, _, tr = run1d(1.0, 1.0, 0.0, s_interior(rho), 10.0, max_evals=2)
    assert tr[1, IDX["rho"]] == pytest.approx(rho, rel=1e-10)
    assert tr[1, IDX["radius"]] == radius
    assert tr[1, IDX["accept"]] == 1 and x[0] == pytest.approx(10.0 - s_interior(rho) * 10.0)


@pytest.mark.parametrize("rho", [-0.05, -1e-3])
def test_step_with_negative_rho_is_rejected(rho):
    x, st, tr = run1d(1.0, 1.0, 0.0, s_interior(rho), 10.0, max_evals=2)
    assert tr[1, IDX["rho"]] == pytest.approx(rho, rel=1e-9)
    assert (tr[1, IDX["accept"]], tr[1, IDX["radius"]], st["rejected"]) == (0, 10.0, 1)
    assert x[0] == 10.0


def test_rejected_step_is_stale():
    # s = 2.5, d = 12: interior p = -30 gives rho = -0.2 (rejected, radius 10, no stop check
    # next time: the iteration count stays); the boundary step of 10 then has
    # rho = (12 - 5) / (30 - 5) = 0.28 (accepted, radius kept)
    x, st, tr = run1d(1.0, 1.0, 0.0, 2.5, 12.0, max_evals=3)
    assert tr[1, IDX["rho"]] == pytest.approx(-0.2, rel=1e-12)
    assert (tr[1, IDX["accept"]], tr[1, IDX["radius"]], tr[1, IDX["iterations"]], tr[1, IDX["rejected"]]) == (0, 10, 1, 1)
    assert tr[1, IDX["sub_iters"]] == 2
    assert tr[2, IDX["rho"]] == pytest.approx(0.28, rel=1e-12)
    assert (tr[2, IDX["accept"]], tr[2, IDX["radius"]], tr[2, IDX["iterations"]]) == (1, 10, 2)
    assert x[0] == pytest.approx(2.0)


def test_doubling_to_the_cap():
    # s = 1, a = b: every boundary step has rho = 1 exactly; the radius doubles to 1000 and
    # stays there; the last step is interior (80 px) and lands on c
    x, st, tr = run1d(1.0, 1.0, 5000.0, 1.0, 0.0)
    assert list(col(tr, "radius")[:10]) == [40, 80, 160, 320, 640, 1000, 1000, 1000, 1000, 1000]
    assert list(col(tr, "sub_iters")[:9]) == [2] * 8 + [1]
    assert np.allclose(np.cumsum([40, 80, 160, 320, 640, 1000, 1000, 1000, 760]), tr[1:10, O.N_TRACE])
    assert x[0] == 5000.0
    # then g = 0: p = 0, no predicted decrease -> stop
    assert st["status"] == 4 and st["evaluations"] == 11


def test_absolute_and_relative_stops_and_iteration_cap():
    # a = 1, b = 2: interior steps halve d = 10: f_k = 50 / 4^k, df_k = 150 / 4^k; rho = 1.5 but
    # the radius never grows (interior steps)
    x, st, tr = run1d(1.0, 2.0, 10.0, 1.0, 0.0)
    assert set(col(tr, "radius")) == {40.0}
    assert np.allclose(col(tr, "rho")[1:], 1.5)
    # |df| < 1e-8 first at k = 17 (150 / 4^17 = 8.7e-9), checked at iteration 18
    assert st["status"] == 1 and st["iterations"] == 18 and st["evaluations"] == 18
    # relative stop with f offset by 1000: 150 / 4^k < 1e-4 (1000 + 50 / 4^(k-1)) first at k = 6
    f = lambda x: 0.5 * (x[0] - 10.0) ** 2 + 1000.0
    g = lambda x: np.array([x[0] - 10.0])
    _, st, _ = O.tr_run(np.array([[[2.0]]]), np.zeros((0, 1, 1)), f, g, [0.0], rel_delta=1e-4)
    assert st["status"] == 2 and st["iterations"] == 7 and st["evaluations"] == 7
    # iteration cap: the 4th stop-strategy call exceeds max_iter = 3
    _, st, _ = run1d(1.0, 2.0, 10.0, 1.0, 0.0, max_iter=3)
    assert st["status"] == 3 and st["iterations"] == 4 and st["evaluations"] == 4


def test_collapse_stops_without_taking_the_step():
    # s = -1: the model gradient points away from c, every step raises f and is rejected; the
    # radius falls by 4 per step, 40 -> 40 / 4^26 = 40 * 2^-52 = 40 * eps: the 26th rejection
    # collapses the radius (radius0 * eps) and the loop stops where x is
    x, st, tr = run1d(1.0, 1.0, 0.0, -1.0, 1.0)
    assert st["status"] == 6 and st["evaluations"] == 27
    assert x[0] == 1.0
    assert tr[-1, IDX["radius"]] == 40.0 * EPS
    assert np.all(col(tr, "accept")[1:] == 0)
    assert list(col(tr, "radius")[1:5]) == [10.0, 2.5, 0.625, 0.15625]
    # the first three rejected steps are the same interior step (|p| = 1 <= radius)
    assert np.all(tr[0:3, O.N_TRACE + 2] == 1.0)


def test_skip_repeat_gives_the_same_path():
    # skip_repeat shrinks the radius past a rejected interior step without evaluating it again
    x1, s1, t1 = run1d(1.0, 1.0, 0.0, -1.0, 1.0)
    x2, s2, t2 = run1d(1.0, 1.0, 0.0, -1.0, 1.0, skip_repeat=1)
    assert x1[0] == x2[0] and s1["status"] == s2["status"] == 6
    assert s2["evaluations"] == s1["evaluations"] - 2
    assert s1["radius_final"] == s2["radius_final"]


def test_zero_gradient_stops_with_no_progress():
    x, st, tr = run1d(1.0, 1.0, 3.0, 1.0, 3.0)
    assert st["status"] == 4 and st["evaluations"] == 2 and x[0] == 3.0


def test_dense_factorisation_gives_the_same_path():
    """dense = 1 factorises the full n x n B as dlib does with the reference's general_matrix;
    on a 3-block model with coupling it takes the same steps as the block-tridiagonal solve."""
    rng = np.random.default_rng(7)
    nb, m = 3, 2
    M = rng.normal(0, 1, (nb * m, nb * m))
    B = M @ M.T + 0.5 * np.eye(nb * m)
    for i in range(nb * m):  # keep B block-tridiagonal
        for j in range(nb * m):
            if abs(i // m - j // m) > 1:
                B[i, j] = 0
    B += 3 * np.eye(nb * m)
    diag = np.array([B[k * m:(k + 1) * m, k * m:(k + 1) * m] for k in range(nb)])
    sub = np.array([B[(k + 1) * m:(k + 2) * m, k * m:(k + 1) * m] for k in range(nb - 1)])
    c = rng.normal(0, 30, nb * m)
    A = B * 0.7
    f = lambda x: 0.5 * (x - c) @ A @ (x - c) + 0.01 * np.sum((x - c) ** 4)
    g = lambda x: A @ (x - c) + 0.04 * (x - c) ** 3
    x1, s1, t1 = O.tr_run(diag, sub, f, g, np.zeros(nb * m), cap=400)
    x2, s2, t2 = O.tr_run(diag, sub, f, g, np.zeros(nb * m), cap=400, dense=1)
    assert s1["evaluations"] == s2["evaluations"] and s1["status"] == s2["status"]
    assert np.allclose(t1[:, :O.N_TRACE], t2[:, :O.N_TRACE], rtol=1e-9, atol=1e-9)
    assert np.max(np.abs(x1 - x2)) <= 1e-9
