"""Pins of the CPU oracle against facts fixed independently of its code.

The reference ships no golden vectors, so the oracle is pinned by closed forms
(scan timing, node map, Sobel on ramps, warp shifts), brute force from the
per-pixel definitions (JtJ, Jtr), an independent interpolation routine
(scipy.ndimage.map_coordinates), finite differences, and trust-region
subproblems with known solutions. Each pin is chosen so that a plausible
mistake (wrong axis, dropped end-of-interval exclusion, sign, off-by-one
interval) fails it; a few tests plant such a mistake to show it is caught.
"""
from fractions import Fraction

import numpy as np
import pytest
from scipy import ndimage, optimize

import oracle as O
from helpers import small_stack


# ---- scan timing and node map (ImageScanParameters.h:7-20; FrameWarper.cpp:223-293) ----

def test_scan_timing_closed_form():
    ref = np.ones((1, 8, 10), np.float32)
    o = O.Oracle().set_reference(ref, 5, line=2.0, interline=1.0)  # interline < line
    s = o.scan()
    assert s["interline"] == 1.01 * 2.0
    assert s["pixel"] == 2.0 / 10
    assert s["frame"] == 8 * 2.02
    assert s["stack"] == s["frame"]
    # 3-D with a large inter-plane gap: capped to frame + nz*frame/(nD-nz)
    ref3 = np.ones((3, 8, 10), np.float32)
    o = O.Oracle().set_reference(ref3, 10, line=1.0, interline=1.0, interframe=100.0)
    s = o.scan()
    cap = 3 * 8.0 / (10 - 3)
    assert s["interframe"] == 8.0 + cap
    assert s["stack"] == 2 * (8.0 + cap) + 8.0


def _exact_u(shape, nD, line, interline, interframe):
    """Node coordinate in exact rationals from the double-valued scan parameters."""
    nz, ny, nx = shape
    L, IL, IF = Fraction(line), Fraction(interline), Fraction(interframe)
    if IL < L:
        IL = Fraction(1.01) * L
    px = L / nx
    fr = ny * IL
    st = (nz - 1) * IF + fr
    cap = nz * fr / (nD - nz)
    if IF - fr > cap:
        IF = fr + cap
        st = (nz - 1) * IF + fr
    out = []
    for z in range(nz):
        for y in range(ny):
            for x in range(nx):
                out.append((z * IF + y * IL + x * px) / st * (nD - 1))
    return out


@pytest.mark.parametrize("shape,nD,line,interline,interframe", [
    ((1, 24, 20), 7, 1.0, 1.0, None),
    ((1, 17, 13), 11, 1.0, 1.3, None),
    ((3, 9, 8), 3 * 4 + 1, 1.0, 1.0, None),
    ((3, 9, 8), 3 * 4 + 1, 1.0, 1.0, 500.0),
])
def test_node_map_closed_form(shape, nD, line, interline, interframe):
    ref = np.ones(shape, np.float32)
    o = O.Oracle().set_reference(ref, nD, line=line, interline=interline, interframe=interframe)
    Di, Df, Dr = o.node_map()
    ny = shape[1]
    IF = interframe if interframe is not None else ny * max(interline, line)
    u = _exact_u(shape, nD, line, interline, IF)
    n = len(u)
    checked = 0
    for p in range(n):
        fl = u[p].numerator // u[p].denominator
        fr = float(u[p] - fl)
        if 1e-9 < fr < 1 - 1e-9:
            assert Di[p] == fl
            assert abs(Df[p] - fr) < 1e-6
            checked += 1
    assert checked > n // 2
    # invariants of D_range (FrameWarper.cpp:269-283)
    assert np.all(np.diff(Di.astype(int)) >= 0)
    assert Di[-1] == nD - 2
    for i in range(nD - 1):
        idx = np.nonzero(Di == i)[0]
        assert Dr[i, 0] == idx[0]
        assert Dr[i, 1] == idx[-1]
    assert Dr[nD - 2, 1] == n - 1
    assert tuple(Dr[nD - 1]) == (0, 0)


def test_node_map_rejects_short_map():
    # more intervals than the last pixel reaches: the reference throws "Something went wrong"
    with pytest.raises(RuntimeError, match="Something went wrong"):
        O.Oracle().set_reference(np.ones((1, 2, 2), np.float32), 20)


# ---- Sobel (Sobel3d.cpp:8-84) ----

def _edge(n):
    w = np.ones(n)
    w[0] = w[-1] = 4.0 / 3.0  # smoothing edge taps 2/3 + 2/3
    return w


def _dedge(n):
    w = np.ones(n)
    w[0] = w[-1] = 0.5  # derivative edge taps -1/2, 1/2
    return w


@pytest.mark.parametrize("shape", [(1, 12, 15), (5, 9, 11)])
@pytest.mark.parametrize("axis", [0, 1, 2])
def test_sobel_ramps(shape, axis):
    nz, ny, nx = shape
    if axis == 0 and nz == 1:
        pytest.skip("no z axis in 2-D")
    alpha, beta = 3.0, 7.0
    zz, yy, xx = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    coord = [zz, yy, xx][axis]
    T = (alpha * coord + beta).astype(np.float32)
    o = O.Oracle().set_reference(T, nz * 3 + 1 if nz > 1 else 5)
    gx, gy, gz = o.gradient()
    ez = _edge(nz) if nz > 1 else np.ones(nz)
    smooth = [ez[:, None, None], _edge(ny)[None, :, None], _edge(nx)[None, None, :]]
    dedge = [_dedge(nz)[:, None, None], _dedge(ny)[None, :, None], _dedge(nx)[None, None, :]]
    g = [gz, gy, gx]
    for a in range(3):
        if a == 0 and nz == 1:
            continue
        if a == axis:
            expect = alpha * dedge[a]
            for b in range(3):
                if b != a:
                    expect = expect * smooth[b]
            np.testing.assert_allclose(g[a], np.broadcast_to(expect, shape), rtol=2e-6)
        else:
            assert not g[a].any(), f"gradient along {a} of a ramp along {axis} must be 0"


def test_sobel_interior_matches_3d_correlation():
    rng = np.random.default_rng(0)
    T = rng.uniform(1, 100, (6, 10, 12)).astype(np.float32)
    o = O.Oracle(fp64=True).set_reference(T, 6 * 2 + 1)
    gx, gy, gz = o.gradient()
    S = np.array([0.25, 0.5, 0.25])
    D = np.array([-0.5, 0.0, 0.5])
    Td = T.astype(np.float64)
    for g, k in ((gx, np.einsum("i,j,k->ijk", S, S, D)), (gy, np.einsum("i,j,k->ijk", S, D, S)),
                 (gz, np.einsum("i,j,k->ijk", D, S, S))):
        ref = ndimage.correlate(Td, k, mode="constant")
        np.testing.assert_allclose(g[1:-1, 1:-1, 1:-1], ref[1:-1, 1:-1, 1:-1], rtol=1e-12, atol=1e-9)


# ---- JtJ and Jtr from the per-pixel definition (FrameWarper.cpp:147-167,296-395) ----

def _rows(o, exclude_end=True):
    """Per-pixel Jacobian rows: node Di gets (1-f) grad, node Di+1 gets f grad; the last
    pixel of each interval contributes nothing (FrameWarper.cpp:309-344)."""
    Di, Df, Dr = o.node_map()
    g = [a.ravel() for a in o.gradient()][:o.n_dim]
    n = o.n_px
    last = np.zeros(n, bool)
    last[-1] = True
    last[:-1] = Di[1:] != Di[:-1]
    w_lo = (1 - Df)
    w_hi = Df
    if exclude_end:
        w_lo = np.where(last, 0, w_lo)
        w_hi = np.where(last, 0, w_hi)
    lo = np.stack([(gc * w_lo).astype(np.float32).astype(np.float64) for gc in g], 1)
    hi = np.stack([(gc * w_hi).astype(np.float32).astype(np.float64) for gc in g], 1)
    return Di.astype(np.int64), lo, hi


def _brute_H(o, exclude_end=True):
    Di, lo, hi = _rows(o, exclude_end)
    m, nD = o.n_dim, o.nD
    H = np.zeros((nD * m, nD * m))
    J = np.zeros((o.n_px, nD * m))
    idx = np.arange(o.n_px)
    for c in range(m):
        J[idx, Di * m + c] = lo[:, c]
        J[idx, (Di + 1) * m + c] = hi[:, c]
    H = J.T @ J
    return H, J


@pytest.mark.parametrize("shape,nD", [((1, 40, 36), 9), ((4, 16, 14), 4 * 3 + 1)])
def test_hessian_brute_force(shape, nD):
    st = small_stack(shape, nD=nD, n_frames=1, seed=4, n_blobs=15)
    o = O.Oracle().set_reference(st["ref_w"], nD)
    H = o.hessian()
    Hb, _ = _brute_H(o)
    np.testing.assert_allclose(H, Hb, rtol=1e-5, atol=1e-6 * np.abs(Hb).max())
    assert np.array_equal(H, H.T)
    assert np.linalg.eigvalsh(H).min() > -1e-9 * np.abs(H).max()
    m = o.n_dim
    for i in range(nD):
        for j in range(nD):
            if abs(i - j) > 1:
                assert not H[i * m:(i + 1) * m, j * m:(j + 1) * m].any()
    # a planted mistake: keeping the interval end pixels changes H beyond the tolerance
    Hbad, _ = _brute_H(o, exclude_end=False)
    assert np.abs(Hbad - H).max() > 1e-3 * np.abs(H).max()


@pytest.mark.parametrize("shape,nD", [((1, 40, 36), 9), ((4, 16, 14), 4 * 3 + 1)])
def test_jacobian_brute_force(shape, nD):
    st = small_stack(shape, nD=nD, n_frames=1, seed=5, n_blobs=15)
    o = O.Oracle().set_reference(st["ref_w"], nD)
    D = st["D_true"][0] + np.random.default_rng(1).normal(0, 0.5, (nD, 3))
    if o.n_dim == 2:
        D[:, 2] = 0
    e, err, wimg = o.get_error(st["frames_w"][0], D, images=True)
    _, jac = o.get_jacobian(st["frames_w"][0], D)
    _, J = _brute_H(o)
    gb = J.T @ err.ravel()
    np.testing.assert_allclose(jac, gb, rtol=1e-5, atol=1e-6 * np.abs(gb).max())
    assert e == pytest.approx(float(np.sum(err.astype(np.float64) ** 2)), rel=1e-6)


def test_jacobian_is_half_gradient_near_alignment():
    """Jtr is the inverse-compositional half gradient of the SSD: with the frame equal to the
    reference and a small displacement it agrees with (1/2) dSSD/dx from central
    differences (a wrong sign, node or axis assignment fails this)."""
    st = small_stack((1, 64, 64), nD=7, n_frames=1, seed=9, n_blobs=40, noise=False, order=1)
    o = O.Oracle(fp64=True).set_reference(st["ref_w"], 7)
    fw = st["ref_w"]
    x0 = np.random.default_rng(0).normal(0, 0.3, 14)
    _, g = o.get_jacobian(fw, o.col2D(x0))
    h = 1e-4
    fd = np.zeros_like(x0)
    for i in range(len(x0)):
        xp, xm = x0.copy(), x0.copy()
        xp[i] += h
        xm[i] -= h
        fd[i] = (o.get_error(fw, o.col2D(xp)) - o.get_error(fw, o.col2D(xm))) / (2 * h)
    cos = g @ (0.5 * fd) / (np.linalg.norm(g) * np.linalg.norm(0.5 * fd))
    assert cos > 0.95, cos
    assert 0.7 < np.linalg.norm(g) / np.linalg.norm(0.5 * fd) < 1.3
    # and Jtr ~ H x for the linearised residual
    H = o.hessian()
    assert (H @ x0) @ g / (np.linalg.norm(H @ x0) * np.linalg.norm(g)) > 0.95


# ---- warp and SSD (FrameWarper.cpp:4-69,397-419) ----

def test_warp_zero_is_identity_and_ssd_zero():
    st = small_stack((1, 30, 28), nD=5, n_frames=1)
    o = O.Oracle().set_reference(st["ref_w"], 5)
    D = np.zeros((5, 3))
    assert np.array_equal(o.warp_interpolated(st["ref_w"], D), st["ref_w"].astype(np.float64))
    assert o.get_error(st["ref_w"], D) == 0.0
    c = np.float32(2.5)
    e = o.get_error(st["ref_w"] + c, D)
    assert e == pytest.approx(float(c) ** 2 * o.n_px, rel=1e-6)


@pytest.mark.parametrize("shape", [(1, 30, 28), (5, 12, 10)])
def test_warp_integer_shift(shape):
    nz = shape[0]
    st = small_stack(shape, nD=nz * 2 + 1 if nz > 1 else 5, n_frames=1)
    nD = st["nD"]
    o = O.Oracle().set_reference(st["ref_w"], nD)
    s = (2, -3, 1 if nz > 1 else 0)
    D = np.tile(np.array(s, float), (nD, 1))
    w = o.warp_interpolated(st["frames_w"][0], D, -9.0)
    expect = np.full(shape, -9.0)
    src = st["frames_w"][0].astype(np.float64)
    zz, yy, xx = np.meshgrid(*[np.arange(n) for n in shape], indexing="ij")
    zs, ys, xs = zz + s[2], yy + s[1], xx + s[0]
    ok = (xs >= 0) & (xs < shape[2]) & (ys >= 0) & (ys < shape[1]) & (zs >= 0) & (zs < shape[0])
    expect[ok] = src[zs[ok], ys[ok], xs[ok]]
    assert np.array_equal(w, expect)
    # shifted frame aligned at the true shift: SSD exactly 0
    fr = np.full(shape, 3.0, np.float32)
    back = (zz - s[2], yy - s[1], xx - s[0])
    okb = (back[2] >= 0) & (back[2] < shape[2]) & (back[1] >= 0) & (back[1] < shape[1]) & \
          (back[0] >= 0) & (back[0] < shape[0])
    fr[okb] = st["ref_w"][back[0][okb], back[1][okb], back[2][okb]]
    assert o.get_error(fr, D) == 0.0


def test_warp_half_pixel_is_neighbour_mean():
    st = small_stack((1, 20, 24), nD=5, n_frames=1)
    o = O.Oracle(fp64=True).set_reference(st["ref_w"], 5)
    D = np.zeros((5, 3))
    D[:, 0] = 0.5
    w = o.warp_interpolated(st["frames_w"][0], D)
    f = st["frames_w"][0].astype(np.float64)
    np.testing.assert_array_equal(w[0, :, :-1], 0.5 * (f[0, :, :-1] + f[0, :, 1:]))
    np.testing.assert_array_equal(w[0, :, -1], 0.5 * f[0, :, -1])


@pytest.mark.parametrize("shape", [(1, 33, 29), (4, 14, 12)])
def test_warp_matches_map_coordinates(shape):
    nz = shape[0]
    nD = nz * 3 + 1 if nz > 1 else 9
    st = small_stack(shape, nD=nD, n_frames=1, seed=8)
    o = O.Oracle(fp64=True).set_reference(st["ref_w"], nD)
    D = np.random.default_rng(2).normal(0, 3.0, (nD, 3))
    if nz == 1:
        D[:, 2] = 0
    w = o.warp_interpolated(st["frames_w"][0], D)
    Di, Df, _ = o.node_map()
    f = Df.reshape(shape)
    i = Di.reshape(shape).astype(int)
    d = [f * D[i + 1, c] + (1 - f) * D[i, c] for c in range(3)]
    zz, yy, xx = np.meshgrid(*[np.arange(n, dtype=float) for n in shape], indexing="ij")
    src = st["frames_w"][0].astype(np.float64)
    if nz == 1:
        ref = ndimage.map_coordinates(src[0], [yy[0] + d[1][0], xx[0] + d[0][0]], order=1,
                                      mode="grid-constant", cval=0.0)[None]
    else:
        ref = ndimage.map_coordinates(src, [zz + d[2], yy + d[1], xx + d[0]], order=1,
                                      mode="grid-constant", cval=0.0)
    np.testing.assert_allclose(w, ref, rtol=1e-10, atol=1e-9)


def test_warp_point_matches_node_interpolation():
    st = small_stack((1, 30, 28), nD=5, n_frames=1)
    o = O.Oracle().set_reference(st["ref_w"], 5)
    D = np.random.default_rng(3).normal(0, 2, (5, 3))
    D[:, 2] = 0
    Di, Df, _ = o.node_map()
    for (x, y) in [(0, 0), (13, 7), (27, 29)]:
        p = y * 28 + x
        f, i = Df[p], int(Di[p])
        np.testing.assert_array_equal(o.warp_point(D, x, y, 0), f * D[i + 1] + (1 - f) * D[i])
    assert not o.warp_point(D, 28, 0, 0).any()  # outside: zero


# ---- trust region (dlib find_min_trust_region / solve_trust_region_subproblem) ----

def _bt_random(nb, m, rng):
    J = rng.normal(size=(4 * nb * m, nb * m))
    # keep only block-tridiagonal couplings: build from local rows
    A = np.zeros((nb * m, nb * m))
    for k in range(nb):
        lo, hi = max(0, k - 1) * m, min(nb, k + 2) * m
        R = rng.normal(size=(3 * m, hi - lo))
        A[lo:hi, lo:hi] += R.T @ R
    A += 1e-3 * np.eye(nb * m)
    for i in range(nb):
        for j in range(nb):
            if abs(i - j) > 1:
                A[i * m:(i + 1) * m, j * m:(j + 1) * m] = 0
    diag = np.stack([A[k * m:(k + 1) * m, k * m:(k + 1) * m] for k in range(nb)])
    sub = np.stack([A[(k + 1) * m:(k + 2) * m, k * m:(k + 1) * m] if k + 1 < nb else np.zeros((m, m))
                    for k in range(nb)])
    return A, diag, sub


@pytest.mark.parametrize("m", [2, 3])
def test_bt_cholesky_solves(m):
    rng = np.random.default_rng(m)
    A, diag, sub = _bt_random(12, m, rng)
    ok, Ld, Ls = O.bt_cholesky(diag, sub, 0.0)
    assert ok
    b = rng.normal(size=12 * m)
    np.testing.assert_allclose(O.bt_solve(Ld, Ls, b), np.linalg.solve(A, b), rtol=1e-9)
    ok, _, _ = O.bt_cholesky(diag, sub, -10 * np.abs(A).max())
    assert not ok


@pytest.mark.parametrize("m", [2, 3])
def test_subproblem_interior_and_boundary(m):
    rng = np.random.default_rng(10 + m)
    A, diag, sub = _bt_random(10, m, rng)
    g = rng.normal(size=10 * m)
    newton = -np.linalg.solve(A, g)
    p, lam, it = O.tr_subproblem_bt(diag, sub, g, 10 * np.linalg.norm(newton))
    assert lam == 0 and it == 1
    np.testing.assert_allclose(p, newton, rtol=1e-9)
    # constrained: the exact solution is p(lam*) with |p(lam*)| = radius
    radius = 0.3 * np.linalg.norm(newton)
    p, lam, it = O.tr_subproblem_bt(diag, sub, g, radius, eps=1e-6, max_iter=100)
    lam_star = optimize.brentq(lambda l: np.linalg.norm(np.linalg.solve(A + l * np.eye(len(g)), -g)) - radius,
                               0, 1e6)
    assert lam == pytest.approx(lam_star, rel=1e-4)
    np.testing.assert_allclose(np.linalg.norm(p), radius, rtol=1e-5)
    # dense and block-tridiagonal storage give the same step
    pd, lamd, itd = O.tr_subproblem_dense(A, g, radius, eps=0.1)
    pb, lamb, itb = O.tr_subproblem_bt(diag, sub, g, radius, eps=0.1)
    np.testing.assert_allclose(pd, pb, rtol=1e-8, atol=1e-12)
    assert itd == itb


def test_align_recovers_displacement_noise_free():
    st = small_stack((1, 96, 96), nD=7, n_frames=2, seed=12, n_blobs=60, noise=False, order=1,
                     sin_amp=2.0, walk_sigma=0.3)
    o = O.Oracle().set_reference(st["ref_w"], 7)
    for i in range(2):
        x, s = o.align(st["frames_w"][i], rel_delta=1e-6, max_iter=50)
        err = np.abs(o.col2D(x) - st["D_true"][i])[:, :2]
        # the generator's d varies within a scan line while a node interval is a line
        # segment; interior nodes are recovered to a small fraction of a pixel
        assert np.median(err) < 0.15, (err, s)
        assert s["f_final"] < o.get_error(st["frames_w"][i], np.zeros((7, 3)))


def test_skip_repeat_keeps_trajectory():
    st = small_stack((1, 64, 64), nD=9, n_frames=3, seed=13)
    o = O.Oracle().set_reference(st["ref_w"], 9)
    for i in range(3):
        x1, s1 = o.align(st["frames_w"][i], rel_delta=1e-6, max_iter=50)
        x2, s2 = o.align(st["frames_w"][i], rel_delta=1e-6, max_iter=50, skip_repeat=1)
        assert np.array_equal(x1, x2)
        assert s1["rejected"] == s2["rejected"] and s1["iterations"] == s2["iterations"]
        assert s2["evaluations"] <= s1["evaluations"]


def test_align_many_matches_align():
    st = small_stack((1, 48, 40), nD=7, n_frames=4, seed=14)
    o = O.Oracle().set_reference(st["ref_w"], 7)
    xs, ss, secs = o.align_many(st["frames_w"], threads=3, rel_delta=1e-4, max_iter=50)
    for i in range(4):
        x, s = o.align(st["frames_w"][i], rel_delta=1e-4, max_iter=50)
        assert np.array_equal(x, xs[i])
    assert np.all(secs > 0)


def test_fp64_mode_close_to_faithful():
    st = small_stack((1, 40, 44), nD=7, n_frames=1, seed=15)
    a = O.Oracle().set_reference(st["ref_w"], 7)
    b = O.Oracle(fp64=True).set_reference(st["ref_w"], 7)
    D = st["D_true"][0]
    ea, ja = a.get_jacobian(st["frames_w"][0], D)
    eb, jb = b.get_jacobian(st["frames_w"][0], D)
    assert ea == pytest.approx(eb, rel=1e-5)
    assert np.linalg.norm(ja - jb) < 1e-4 * np.linalg.norm(jb)


# ---- forward splat (FrameWarper.cpp:71-136) ----

@pytest.mark.parametrize("shape", [(1, 26, 30), (4, 12, 14)])
def test_splat_identity_and_mass(shape):
    nz = shape[0]
    nD = nz * 2 + 1 if nz > 1 else 5
    st = small_stack(shape, nD=nD, n_frames=1, seed=17)
    o = O.Oracle(fp64=True).set_reference(st["ref_w"], nD)
    f = st["frames_w"][0].astype(np.float64)
    w, c = o.warp_image(st["frames_w"][0], np.zeros((nD, 3)))
    assert np.array_equal(w, f) and np.array_equal(c, np.ones(shape))
    # any displacement: each source that lands spreads total weight 1 (mass is conserved)
    D = np.random.default_rng(4).normal(0, 2.0, (nD, 3))
    if nz == 1:
        D[:, 2] = 0
    w, c = o.warp_image(st["frames_w"][0], D)
    Di, Df, _ = o.node_map()
    fr = Df.reshape(shape)
    i = Di.reshape(shape).astype(int)
    d = [fr * D[i + 1, k] + (1 - fr) * D[i, k] for k in range(3)]
    zz, yy, xx = np.meshgrid(*[np.arange(n, dtype=float) for n in shape], indexing="ij")
    l0 = [np.floor(xx - d[0]), np.floor(yy - d[1]), np.floor(zz - d[2])]
    ok = (l0[0] >= 0) & (l0[0] <= shape[2] - 1) & (l0[1] >= 0) & (l0[1] <= shape[1] - 1)
    if nz > 1:
        ok &= (l0[2] >= 0) & (l0[2] <= nz - 1)
    assert c.sum() == pytest.approx(ok.sum(), rel=1e-12)
    assert w.sum() == pytest.approx(f[ok].sum(), rel=1e-12)


def test_splat_integer_and_half_shift():
    st = small_stack((1, 20, 22), nD=5, n_frames=1, seed=18)
    o = O.Oracle(fp64=True).set_reference(st["ref_w"], 5)
    f = st["frames_w"][0].astype(np.float64)[0]
    D = np.zeros((5, 3))
    D[:, 0], D[:, 1] = 2, -1  # source (x, y) lands on (x - 2, y + 1)
    w, c = o.warp_image(st["frames_w"][0], D)
    expect = np.zeros((20, 22))
    expect[1:, :20] = f[:19, 2:]
    assert np.array_equal(w[0], expect)
    assert np.array_equal(c[0], (expect != 0).astype(float))
    D = np.zeros((5, 3))
    D[:, 0] = -0.5  # each source splits evenly between x + 0 and x + 1 (last column keeps both)
    w, c = o.warp_image(st["frames_w"][0], D)
    expect = 0.5 * f.copy()
    expect[:, 1:] += 0.5 * f[:, :-1]
    expect[:, -1] += 0.5 * f[:, -1]
    np.testing.assert_allclose(w[0], expect, rtol=1e-15)
