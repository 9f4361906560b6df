"""Parity of the HIP path (through the C ABI) against the CPU oracle on seeded inputs.

Tolerances (the path's parity contract): SSD within 1e-4 relative, JtJ / Jtr within 1e-3
relative, converged node displacements within 0.02 px. Index work (node map,
interval ranges) and the Sobel gradient, which the reference computes in plain
float with a fixed operation order, are compared bit for bit.
"""
import numpy as np
import pytest

import oracle as O
import fa_amd as F
from helpers import LONG_CHAIN_CASE, MANY_NODE_CASES, PARITY_CASES, rel_err, small_stack, truth_start

pytestmark = pytest.mark.gpu

ERR_TOL = 1e-4
JAC_TOL = 1e-3
D_TOL = 0.02


@pytest.fixture(scope="module")
def ctx():
    c = F.Context(0)
    yield c
    c.close()


CASES = [c[:4] for c in PARITY_CASES]
GEN = {c[0]: c[4] for c in PARITY_CASES + MANY_NODE_CASES}

# Widths that are not a multiple of 4 stage their rows sample by sample (the other widths in
# groups of four); 3-D with an odd width. Frames 768 or more wide (a multiple of 256) are cut
# into strip work items of 256 columns.
ODD_WIDTH_CASES = [
    ("2d_width70", (1, 64, 70), 9, dict(line=1.0, interline=1.0), {}),
    ("3d_width37", (4, 40, 37), 4 * 4 + 1, dict(line=1.0, interline=1.0), dict(n_blobs=90, sigma=(1.5, 4.0))),
]
STRIP_CASES = [
    ("2d_strips", (1, 48, 768), 9, dict(line=1.0, interline=1.0), dict(n_blobs=220)),
    ("3d_strips", (3, 20, 768), 3 * 4 + 1, dict(line=1.0, interline=1.0), dict(n_blobs=260, sigma=(1.5, 4.0))),
]
GEN.update({c[0]: c[4] for c in STRIP_CASES})


def _pair(ctx, shape, nD, scan, seed=3, name=None, **kw):
    kw = {**GEN.get(name, {}), **kw}
    st = small_stack(shape, nD=nD, seed=seed, **kw)
    orc = O.Oracle().set_reference(st["ref_w"], nD, **scan)
    ref = F.Reference(ctx, st["ref_w"], nD, **scan)
    return st, orc, ref


@pytest.mark.parametrize("name,shape,nD,scan", CASES)
def test_node_map_bit_exact(ctx, name, shape, nD, scan):
    st, orc, ref = _pair(ctx, shape, nD, scan, name=name)
    Di_o, Df_o, Dr_o = orc.node_map()
    Di_g, Df_g, Dr_g = ref.node_map()
    assert np.array_equal(Di_o, Di_g)
    assert np.array_equal(Df_o.astype(np.float32), Df_g)
    assert np.array_equal(Dr_o, Dr_g)


@pytest.mark.parametrize("name,shape,nD,scan", CASES)
def test_sobel_bit_exact(ctx, name, shape, nD, scan):
    st, orc, ref = _pair(ctx, shape, nD, scan, name=name)
    go = orc.gradient()
    gg = ref.gradient()
    for c in range(ref.n_dim):
        assert np.array_equal(go[c].astype(np.float32), gg[c]), f"component {c}"


@pytest.mark.parametrize("name,shape,nD,scan", CASES + [c[:4] for c in STRIP_CASES])
def test_hessian(ctx, name, shape, nD, scan):
    st, orc, ref = _pair(ctx, shape, nD, scan, name=name)
    Ho, Hg = orc.hessian(), ref.H
    assert rel_err(Hg, Ho) < JAC_TOL
    # block-tridiagonal structure: zero beyond the first off-diagonal block
    m = ref.n_dim
    for i in range(nD):
        for j in range(nD):
            if abs(i - j) > 1:
                assert not Hg[i * m:(i + 1) * m, j * m:(j + 1) * m].any()


def _displacements(st, nD, rng):
    yield np.zeros((nD, 3))
    yield st["D_true"][0]
    yield st["D_true"][0] + rng.normal(0, 0.7, (nD, 3))
    yield rng.normal(0, 4.0, (nD, 3))


@pytest.mark.parametrize("name,shape,nD,scan", CASES)
@pytest.mark.parametrize("dtype", ["float32", "uint16"])
def test_error_and_jacobian(ctx, name, shape, nD, scan, dtype):
    st, orc, ref = _pair(ctx, shape, nD, scan, name=name, dtype=dtype)
    rng = np.random.default_rng(11)
    for f_i in range(len(st["frames"])):
        fr = ref.register(st["frames"][f_i], offset=1.0)  # the aligner's +1 on the device
        fw = st["frames_w"][f_i]
        for D in _displacements(st, nD, rng):
            if ref.n_dim == 2:
                D = D.copy()
                D[:, 2] = 0
            e_o, j_o = orc.get_jacobian(fw, D)
            e_g, j_g = fr.jacobian(D)
            assert abs(e_g - e_o) <= ERR_TOL * abs(e_o), (e_g, e_o)
            assert abs(fr.error(D) - e_g) <= 1e-12 * abs(e_g)
            assert rel_err(j_g, j_o) < JAC_TOL, rel_err(j_g, j_o)
        fr.release()




@pytest.mark.parametrize("name,shape,nD,scan,gen", ODD_WIDTH_CASES + STRIP_CASES)
@pytest.mark.parametrize("dtype", ["float32", "uint16"])
def test_error_and_jacobian_odd_width(ctx, name, shape, nD, scan, gen, dtype):
    st, orc, ref = _pair(ctx, shape, nD, scan, dtype=dtype, **gen)
    rng = np.random.default_rng(12)
    for f_i in range(len(st["frames"])):
        fr = ref.register(st["frames"][f_i], offset=1.0)
        fw = st["frames_w"][f_i]
        for D in _displacements(st, nD, rng):
            if ref.n_dim == 2:
                D = D.copy()
                D[:, 2] = 0
            e_o, j_o = orc.get_jacobian(fw, D)
            e_g, j_g = fr.jacobian(D)
            assert abs(e_g - e_o) <= ERR_TOL * abs(e_o), (e_g, e_o)
            assert rel_err(j_g, j_o) < JAC_TOL, rel_err(j_g, j_o)
        fr.release()


@pytest.mark.parametrize("name,shape,nD,scan,gen",
                         [c[:4] + (GEN.get(c[0], {}),) for c in CASES] + ODD_WIDTH_CASES + STRIP_CASES
                         + [LONG_CHAIN_CASE])
def test_staged_and_global_sampling_agree(ctx, name, shape, nD, scan, gen):
    """Work items sample the frame from their LDS tile or, when it does not fit, from global
    memory with the same arithmetic: a batch run with staging switched off, or with an LDS
    budget that stages only some items, gives the same bits."""
    st, orc, ref = _pair(ctx, shape, nD, scan, dtype="uint16", **gen)
    kw = dict(rel_delta=1e-4, max_iter=50, skip_repeat=1)
    frames = [ref.register(f, offset=1.0) for f in st["frames"]]
    batch = F.Batch(frames, **kw)
    x1, s1, _ = batch.run()
    full = batch.tuning["lds_floats"]
    for tune in (dict(use_lds=0), dict(use_lds=1, lds_floats=max(16, full * 2 // 3))):
        batch.set_tuning(**tune)
        x0, s0, _ = batch.run()
        for a, b, sa, sb in zip(x1, x0, s1, s0):
            assert np.array_equal(a, b), tune
            assert sa["evaluations"] == sb["evaluations"]
    batch.close()


@pytest.mark.parametrize("name,shape,nD,scan", CASES + [c[:4] for c in STRIP_CASES])
def test_warp_interpolated(ctx, name, shape, nD, scan):
    st, orc, ref = _pair(ctx, shape, nD, scan, name=name)
    fr = ref.register(st["frames_w"][0])
    rng = np.random.default_rng(5)
    for D in _displacements(st, nD, rng):
        for invalid in (0.0, -1.0):
            wo = orc.warp_interpolated(st["frames_w"][0], D, invalid)
            wg = fr.warp_interpolated(D, invalid)
            assert np.array_equal(wo == invalid, wg == invalid)
            assert np.max(np.abs(wg - wo)) <= 1e-5 * np.max(np.abs(wo))


def test_zero_and_integer_shift_exact(ctx):
    st, orc, ref = _pair(ctx, (1, 64, 64), 9, dict(line=1.0, interline=1.0))
    fr = ref.register(st["ref_w"])
    assert fr.error(np.zeros((9, 3))) == 0.0
    D = np.zeros((9, 3))
    D[:, 0] = 2
    D[:, 1] = -3
    w = fr.warp_interpolated(D, -7.0)
    expect = np.full((1, 64, 64), -7.0, np.float32)
    expect[0, 3:, :62] = st["ref_w"][0, :61, 2:]
    assert np.array_equal(w, expect)


@pytest.mark.parametrize("name,shape,nD,scan", CASES + [c[:4] for c in STRIP_CASES])
def test_align_matches_oracle(ctx, name, shape, nD, scan):
    st, orc, ref = _pair(ctx, shape, nD, scan, name=name)
    kw = dict(rel_delta=1e-4, max_iter=50, skip_repeat=1)
    for f_i in range(len(st["frames"])):
        fw = st["frames_w"][f_i]
        x_o, s_o = orc.align(fw, **kw)
        fr = ref.register(fw)
        x_g, s_g = fr.align(**kw)
        assert np.max(np.abs(x_g - x_o)) <= D_TOL, (np.max(np.abs(x_g - x_o)), s_o, s_g)
        # the relative stop rule fixes the final SSD only to ~rel_delta, so the SSDs are
        # compared at the same point: the GPU's result
        e_o = orc.get_error(fw, orc.col2D(x_g))
        assert abs(s_g["f_final"] - e_o) <= ERR_TOL * e_o
        assert abs(s_g["f_final"] - s_o["f_final"]) <= 10 * kw["rel_delta"] * s_o["f_final"]
        fr.release()


@pytest.mark.parametrize("name,shape,nD,scan", [c[:4] for c in MANY_NODE_CASES])
def test_align_many_nodes_matches_oracle(ctx, name, shape, nD, scan):
    """More node blocks than lanes per level of the cyclic reduction (150 and 161 nodes)."""
    st, orc, ref = _pair(ctx, shape, nD, scan, name=name)
    assert rel_err(ref.H, orc.hessian()) < JAC_TOL
    kw = dict(rel_delta=1e-4, max_iter=50, skip_repeat=1)
    frames = [ref.register(fw) for fw in st["frames_w"]]
    x0 = [truth_start(st, f, ref.n_dim) for f in range(len(frames))]
    batch = F.Batch(frames, **kw)
    xs, stats, _ = batch.run(x0)
    for f_i, fw in enumerate(st["frames_w"]):
        x_o, s_o = orc.align(fw, x0=x0[f_i], **kw)
        assert np.max(np.abs(xs[f_i] - x_o)) <= D_TOL, (np.max(np.abs(xs[f_i] - x_o)), s_o, stats[f_i])
        x_g, s_g = frames[f_i].align(x0=x0[f_i], **kw)
        assert np.array_equal(x_g, xs[f_i])


def test_batch_same_alone_and_in_mix_update_widths(ctx):
    """Frames whose update runs with 256 threads (150 nodes) beside frames whose update runs
    with one wave (9 nodes) in one batch: each result is bit-identical to the frame alone, and
    the long chain's result agrees with the oracle."""
    name, shape, nD, scan, _ = MANY_NODE_CASES[0]
    a, orc_a, ref_a = _pair(ctx, shape, nD, scan, name=name)
    b, orc_b, ref_b = _pair(ctx, (1, 96, 80), 9, dict(line=1.0, interline=1.0), seed=21)
    kw = dict(rel_delta=1e-4, max_iter=50, skip_repeat=1)
    fa_ = [ref_a.register(f) for f in a["frames_w"]]
    fb_ = [ref_b.register(f) for f in b["frames_w"]]
    x0 = [truth_start(a, i, 2) for i in range(len(fa_))] + [None] * len(fb_)
    frames = fa_ + fb_
    alone = [f.align(x0=x, **kw) for f, x in zip(frames, x0)]
    xs, stats, _ = F.Batch(frames, **kw).run([x if x is not None else np.zeros(f.ref.nD * f.ref.n_dim)
                                              for f, x in zip(frames, x0)])
    for (x1, s1), x2, s2 in zip(alone, xs, stats):
        assert np.array_equal(x1, x2)
        assert s1["evaluations"] == s2["evaluations"]
    x_o, _ = orc_a.align(a["frames_w"][0], x0=x0[0], **kw)
    assert np.max(np.abs(xs[0] - x_o)) <= D_TOL


def test_batch_update_classes_launched_apart(ctx):
    """A batch with more frames than can hold their update workgroups at once runs its update
    classes one launch after another (a batch small enough runs them in one launch): 256-thread
    3-D frames (161 nodes) beside a thousand-odd one-wave 2-D frames, each result bit-identical
    to its frame aligned alone."""
    name, shape, nD, scan, _ = MANY_NODE_CASES[1]
    a, _, ref_a = _pair(ctx, shape, nD, scan, name=name)
    b, _, ref_b = _pair(ctx, (1, 96, 80), 9, dict(line=1.0, interline=1.0), seed=21)
    info = ctx.info()
    # the 3-D frames' update needs at least its factor blocks and solve vector in LDS
    # (3 nD 9 + 3 nD doubles), so at most this many workgroups are resident at once; the batch
    # runs as two parts of about half its pixels each, so the part with the 3-D frames holds
    # more than that when the 2-D frames number twice as many and a margin
    resident = info["n_cu"] * max(1, info["lds_per_block"] // (8 * 30 * nD))
    reps = (2 * resident + 64) // len(b["frames_w"]) + 1
    kw = dict(rel_delta=1e-4, max_iter=50, skip_repeat=1)
    fa_ = [ref_a.register(f) for f in a["frames_w"]]
    fb_ = [ref_b.register(b["frames_w"][i % len(b["frames_w"])]) for i in range(reps * len(b["frames_w"]))]
    assert len(fb_) > 2 * resident + 32
    alone_a = [f.align(**kw) for f in fa_]
    alone_b = [f.align(**kw) for f in fb_[:len(b["frames_w"])]]
    xs, stats, _ = F.Batch(fa_ + fb_, **kw).run()
    for i, (x1, s1) in enumerate(alone_a):
        assert np.array_equal(x1, xs[i])
        assert s1["evaluations"] == stats[i]["evaluations"]
    for j in range(len(fb_)):
        x1, s1 = alone_b[j % len(b["frames_w"])]
        assert np.array_equal(x1, xs[len(fa_) + j])
        assert s1["evaluations"] == stats[len(fa_) + j]["evaluations"]
    for f in fa_ + fb_:
        f.release()


def test_batch_same_alone_and_in_mix(ctx):
    """Each frame's result is bit-identical alone and inside a mixed batch (two references of
    different size, uint16 and float frames)."""
    a, orc_a, ref_a = _pair(ctx, (1, 96, 80), 9, dict(line=1.0, interline=1.0), seed=21)
    b, orc_b, ref_b = _pair(ctx, (1, 64, 72), 13, dict(line=1.0, interline=1.3), seed=22,
                            dtype="uint16")
    kw = dict(rel_delta=1e-4, max_iter=50, skip_repeat=1)
    frames = [ref_a.register(f) for f in a["frames_w"]] + [ref_b.register(f) for f in b["frames"]]
    alone = [f.align(**kw) for f in frames]
    batch = F.Batch(frames, **kw)
    xs, stats, _ = batch.run()
    for (x1, s1), x2, s2 in zip(alone, xs, stats):
        assert np.array_equal(x1, x2)
        assert s1["evaluations"] == s2["evaluations"]
    # and against the oracle
    for i, fw in enumerate(a["frames_w"]):
        x_o, _ = orc_a.align(fw, **kw)
        assert np.max(np.abs(xs[i] - x_o)) <= D_TOL
    xs2, _, _ = batch.run()
    for x1, x2 in zip(xs, xs2):
        assert np.array_equal(x1, x2)


def test_errors_are_reported(ctx):
    st, orc, ref = _pair(ctx, (1, 64, 64), 9, dict(line=1.0, interline=1.0))
    with pytest.raises(F.FaError):
        ref.register(np.zeros((10, 10), np.float32))
    with pytest.raises(F.FaError):
        F.Reference(ctx, st["ref_w"], 1)
    with pytest.raises(F.FaError):
        F.Reference(ctx, np.zeros((1, 1, 8), np.float32), 4)
    fr = ref.register(st["frames_w"][0])
    # displacement far outside: every sample missing, SSD 0, Jtr 0
    D = np.full((9, 3), 1e7)
    e, j = fr.jacobian(D)
    assert e == 0.0 and not j.any()
    D = np.full((9, 3), np.nan)
    e, j = fr.jacobian(D)
    assert e == 0.0 and not j.any()


@pytest.mark.parametrize("per_ref", [36, 75])
def test_split_batch_same_as_alone(ctx, per_ref):
    """A batch of >= 64 frames runs as two parts (>= 144 frames: three) on their own streams
    with interleaved rounds; every frame's result is bit-identical to aligning it alone."""
    a, _, ref_a = _pair(ctx, (1, 48, 40), 7, dict(line=1.0, interline=1.0), seed=41, n_frames=per_ref)
    b, _, ref_b = _pair(ctx, (1, 40, 56), 9, dict(line=1.0, interline=1.0), seed=42, n_frames=per_ref,
                        dtype="uint16")
    kw = dict(rel_delta=1e-4, max_iter=50, skip_repeat=1)
    frames = [ref_a.register(f) for f in a["frames_w"]] + [ref_b.register(f) for f in b["frames"]]
    batch = F.Batch(frames, **kw)
    batch.set_timing(True)
    xs, stats, timing = batch.run()
    assert timing["pass_launches"] > 0 and timing["pass_bytes"] > 0
    n = len(frames)
    for i in sorted({0, n // 3 - 1, n // 3, n // 2 - 1, n // 2, 2 * n // 3, n - 1}):
        x1, s1 = frames[i].align(**kw)
        assert np.array_equal(x1, xs[i]), i
        assert s1["evaluations"] == stats[i]["evaluations"]
    D = [np.zeros((f.ref.nD, 3)) for f in frames]
    err, jac = batch.eval(D)
    for i in (3, 40):
        e, j = frames[i].jacobian(D[i])
        assert e == err[i] and np.array_equal(j, jac[i])


@pytest.mark.parametrize("name,shape,nD,scan", CASES)
def test_warp_image_splat(ctx, name, shape, nD, scan):
    st, orc, ref = _pair(ctx, shape, nD, scan, name=name, dtype="uint16")
    fr = ref.register(st["frames"][0])  # uint16 counts, +1 registered for alignment
    raw = st["frames"][0].astype(np.float32)
    w0, c0 = fr.warp_image(np.zeros((nD, 3)))
    assert np.array_equal(w0, raw) and np.array_equal(c0, np.ones(shape, np.float32))
    rng = np.random.default_rng(9)
    for D in (st["D_true"][0], rng.normal(0, 3.0, (nD, 3))):
        if ref.n_dim == 2:
            D = D.copy()
            D[:, 2] = 0
        wo, co = orc.warp_image(raw, D)
        wg, cg = fr.warp_image(D)
        assert np.max(np.abs(wg - wo)) <= 1e-5 * max(1.0, np.max(np.abs(wo)))
        assert np.max(np.abs(cg - co)) <= 1e-5
        w1, _ = fr.warp_image(D, apply_offset=True)  # + offset x coverage, up to float rounding
        assert np.max(np.abs((w1 - wg) - cg)) <= 1e-5 * max(1.0, np.max(np.abs(w1)))


def test_graph_replays_match_eager(ctx):
    """Trust-region rounds replayed as captured hipGraphs give the same bits as eager launches,
    on every replay, with eager work and synchronisations in between."""
    a, _, ref_a = _pair(ctx, (1, 64, 56), 9, dict(line=1.0, interline=1.0), seed=51, n_frames=6)
    kw = dict(rel_delta=1e-4, max_iter=50, skip_repeat=1)
    frames = [ref_a.register(f) for f in a["frames_w"]]
    batch = F.Batch(frames, **kw)
    batch.set_tuning(use_graph=0)
    x_eager, s_eager, _ = batch.run()
    batch.set_tuning(use_graph=1, check_every=3)
    for rep in range(4):
        xs, stats, timing = batch.run()
        for x1, x2, s1, s2 in zip(x_eager, xs, s_eager, stats):
            assert np.array_equal(x1, x2), rep
            assert s1["evaluations"] == s2["evaluations"]
        e, _ = frames[rep].jacobian(np.zeros((9, 3)))  # eager work between replays
        assert e > 0
    batch.set_tuning(use_graph=0)
    xs, _, _ = batch.run()
    assert all(np.array_equal(x1, x2) for x1, x2 in zip(x_eager, xs))
