"""Device front end of the aligner (fa_aligner_*, FrameWarpAligner::setReference/addFrame)
against the front-end oracle (oracle/frontend.py) and the warper oracle: rigid phase-correlation
start, blur + 1, start selection and the trust region from the chosen start."""
import numpy as np
import pytest

import oracle as O
import frontend as FE
import fa_amd as F
from fa_amd import synth
from test_frontend_oracle import _phantom_pair

pytestmark = pytest.mark.gpu

D_TOL = 0.02
ERR_TOL = 1e-4


@pytest.fixture(scope="module")
def ctx():
    c = F.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("shape,shift,binning", [((1, 64, 80), (5, -3, 0), 1), ((1, 64, 64), (-7, 4, 0), 1),
                                                 ((6, 48, 40), (3, -2, 1), 1), ((1, 64, 64), (6, -4, 0), 2)])
def test_rigid_shift_matches_oracle(ctx, shape, shift, binning):
    ref, frame = _phantom_pair(shape, shift)
    n_dim = 3 if shape[0] > 1 else 2
    al = F.Aligner(ctx, ref, n_resampling_points=4, spatial_binning=binning)
    pc = FE.PhaseCorrelator(FE.downsample(ref, binning).shape).set_reference(FE.downsample(ref, binning))
    s_o = FE.rigid_shift(pc, frame, binning)
    s_g = al.rigid_shift(frame)
    assert s_g == s_o
    assert s_g[:n_dim] == tuple(float(v) for v in shift[:n_dim])
    # uint16 input gives the same shift as its float conversion
    q = np.clip(np.rint(frame), 0, 65535).astype(np.uint16)
    assert al.rigid_shift(q) == FE.rigid_shift(pc, q.astype(np.float32), binning)
    al.close()


@pytest.mark.parametrize("smoothing", [0.0, 1.5, 3.0])
@pytest.mark.parametrize("dtype", ["float32", "uint16"])
def test_preprocess_matches_oracle(ctx, smoothing, dtype):
    st = synth.make_stack((3, 40, 52), 1, 7, seed=4, n_blobs=50, sigma=(1.5, 4.0), dtype=dtype)
    al = F.Aligner(ctx, st["reference"], n_resampling_points=2, smoothing=smoothing)
    raw = st["frames"][0]
    out = al.preprocess(raw)
    expect = FE.smooth_stack(raw.astype(np.float32), smoothing) + np.float32(1.0)
    assert np.max(np.abs(out - expect)) <= 1e-5 * np.max(np.abs(expect))
    if smoothing == 0:
        assert np.array_equal(out, expect)
    al.close()


def _offset_frames(shape, nD, n, offset, seed, n_blobs=160, scale=4.0):
    """Frames of the generator's motion model (gentle: random walk 0.1 px per node, 0.5 px
    sinusoid) plus a rigid offset (x, y) on every node, so that the rigid start is the better
    one for the offset frames; raw float photon counts (phantom scaled by `scale`)."""
    rng = np.random.default_rng(seed)
    n_dim = 3 if shape[0] > 1 else 2
    ph = scale * synth.blob_phantom(shape, n_blobs, (2.0, 5.0), rng=rng)
    u = synth.node_coordinate(shape, nD)
    frames, Ds = [], []
    for i in range(n):
        D = synth.node_displacements(nD, n_dim, rng, walk_sigma=0.1, sin_amp=0.5)
        D[:, :2] += offset[i]
        clean = np.maximum(synth.displaced(ph, synth.displacement_field(u, D)), 0)
        frames.append(rng.poisson(clean).astype(np.float32))
        Ds.append(D)
    return rng.poisson(ph).astype(np.float32), frames, Ds


@pytest.mark.parametrize("smoothing", [0.0, 1.0])
def test_add_frames_matches_oracle(ctx, smoothing):
    shape, n_res = (1, 128, 112), 8
    nD = n_res + 1
    ref, frames, _ = _offset_frames(shape, nD, 4, [(0, 0), (3, -2), (-2, 3), (0.4, 0.3)], seed=11)
    al = F.Aligner(ctx, ref, n_resampling_points=n_res, smoothing=smoothing)
    kw = dict(rel_delta=1e-4, max_iter=50, skip_repeat=1)
    D_g, res, imgs = al.add_frames(frames, outputs=("warped",), **kw)
    realigned = imgs["warped"]
    ref_w = FE.smooth_stack(ref, smoothing) + np.float32(1.0)
    orc = O.Oracle().set_reference(ref_w, nD)
    pc = FE.PhaseCorrelator(shape).set_reference(ref)
    rng = np.random.default_rng(0)
    starts, compared = [], []
    for i, raw in enumerate(frames):
        fw = FE.smooth_stack(raw, smoothing) + np.float32(1.0)
        s = FE.rigid_shift(pc, raw, 1)
        assert res[i]["rigid_shift"] == s
        Dr = np.tile(np.array(s), (nD, 1))
        e0, er = orc.get_error(fw, np.zeros((nD, 3))), orc.get_error(fw, Dr)
        assert abs(res[i]["start_error"][0] - e0) <= ERR_TOL * e0
        assert abs(res[i]["start_error"][2] - er) <= ERR_TOL * er
        k = FE.choose_start([e0, e0, er])
        assert res[i]["start"] == k
        starts.append(k)
        w_o = orc.warp_interpolated(fw, D_g[i], 0.0)
        assert np.max(np.abs(realigned[i] - w_o)) <= 1e-5 * np.max(np.abs(w_o))
        # the displacement is compared where the trust region from this start is well posed:
        # a 1e-5 px perturbation of the start moves the oracle's result by < 0.005 px
        x0 = Dr[:, :2].reshape(-1) if k == 2 else np.zeros(2 * nD)
        x_o, _ = orc.align(fw, x0=x0, **kw)
        moves = [np.max(np.abs(orc.align(fw, x0=x0 + rng.normal(0, 1e-5, x0.size), **kw)[0] - x_o))
                 for _ in range(3)]
        if max(moves) < 0.005:
            assert np.max(np.abs(D_g[i][:, :2].reshape(-1) - x_o)) <= D_TOL
            compared.append(k)
    assert 2 in starts and 0 in starts   # both kinds of start occur in this set
    # From the integer rigid start the reference's trust region (radius 40) is seldom well posed
    # on this workload (1e-5 px start perturbations move the oracle's result by 0.1-1 px), so
    # the displacement parity is certified on the zero-start frames here; the rigid start itself
    # is compared bit for bit above, and the trust region from a given start is pinned by
    # test_gpu_parity.test_align_many_nodes_matches_oracle.
    assert 0 in compared


def test_translation_type_returns_the_start(ctx):
    shape = (1, 64, 64)
    ref, frames, _ = _offset_frames(shape, 9, 2, [(5, -3), (0, 0)], seed=12)
    al = F.Aligner(ctx, ref, type="translation", n_resampling_points=8)
    D, res, _ = al.add_frames(frames)
    for Di, r in zip(D, res):
        expect = np.tile(np.array(r["rigid_shift"])[:2], (al.nD, 1)) if r["start"] == 2 else np.zeros((al.nD, 2))
        assert np.array_equal(Di[:, :2], expect) and not Di[:, 2].any()
        assert r["stats"]["evaluations"] == 0
    al.close()


def test_add_frames_kept_state_between_calls(ctx):
    """The arrays, frames and batch add_frames keeps between calls: calls of the same and of
    other frame counts, with other frames and stop rules, and after clear_temp, each give what
    a fresh aligner gives for those frames."""
    st = synth.make_stack((1, 96, 80), 7, 9, seed=14, n_blobs=120, dtype="uint16")
    fr = list(st["frames"])
    kw = dict(rel_delta=1e-4, max_iter=50, skip_repeat=1)
    calls = [(fr[:3], kw), (fr[3:6], kw), (fr[:3], dict(rel_delta=1e-3, max_iter=20, skip_repeat=1)),
             (fr[1:6], kw), (fr[:2], kw)]
    al = F.Aligner(ctx, st["reference"], n_resampling_points=8, smoothing=1.0)
    for k, (frames, p) in enumerate(calls):
        if k == 3:
            al.clear_temp()
        D1, r1, im1 = al.add_frames(frames, outputs=("realigned",), **p)
        fresh = F.Aligner(ctx, st["reference"], n_resampling_points=8, smoothing=1.0)
        D2, r2, im2 = fresh.add_frames(frames, outputs=("realigned",), **p)
        fresh.close()
        for a, b in zip(D1, D2):
            assert np.array_equal(a, b), k
        for a, b in zip(r1, r2):
            assert a["stats"] == b["stats"] and a["start"] == b["start"], k
            assert a["correlation"] == b["correlation"] and a["coverage"] == b["coverage"], k
        assert np.array_equal(im1["realigned"], im2["realigned"]), k
    al.close()


@pytest.mark.parametrize("shape,n_res", [((1, 96, 80), 8), ((4, 40, 36), 5)])
def test_add_frames_outputs_and_metrics(ctx, shape, n_res):
    """RealignmentResult's images and metrics (FrameWarpAligner.cpp:218-240,317-346) at the
    device's final D, against the oracle's warps of the same raw frame at that D."""
    nD = n_res * shape[0] + 1
    st = synth.make_stack(shape, 3, nD, seed=8, n_blobs=120, sigma=(1.5, 4.0), dtype="uint16")
    al = F.Aligner(ctx, st["reference"], n_resampling_points=n_res, smoothing=1.0)
    D_g, res, imgs = al.add_frames(list(st["frames"]), outputs=("realigned", "preserving", "mask", "warped"),
                                   rel_delta=1e-4, max_iter=50, skip_repeat=1)
    ref_w = FE.smooth_stack(st["reference"], 1.0) + np.float32(1.0)
    orc = O.Oracle().set_reference(ref_w, nD)
    for i, raw in enumerate(st["frames"]):
        rawf = raw.astype(np.float32)
        wimg, cov = orc.warp_image(rawf, D_g[i])
        assert np.max(np.abs(imgs["mask"][i] - cov)) <= 1e-5
        assert res[i]["coverage"] == np.count_nonzero(cov) / cov.size
        pres = FE.preserving(wimg, cov).astype(np.int64)
        assert np.max(np.abs(imgs["preserving"][i].astype(np.int64) - pres)) <= 1
        realg = FE.to_u16(orc.warp_interpolated(rawf, D_g[i], 0.0)).astype(np.int64)
        diff = np.abs(imgs["realigned"][i].astype(np.int64) - realg)
        assert diff.max() <= 1 and np.count_nonzero(diff) <= 1e-3 * diff.size   # only at .5 ties
        fw = FE.smooth_stack(raw, 1.0) + np.float32(1.0)
        warped_o = orc.warp_interpolated(fw, D_g[i], 0.0)
        assert np.max(np.abs(imgs["warped"][i] - warped_o)) <= 1e-5 * np.max(np.abs(warped_o))
        assert abs(res[i]["correlation"] - FE.correlation(warped_o, ref_w, cov)) <= 1e-5
        assert abs(res[i]["unaligned_correlation"] - FE.correlation(fw, ref_w, cov)) <= 1e-5
    al.close()


@pytest.mark.parametrize("shape,n_res,dtype", [((1, 96, 80), 8, "uint16"), ((4, 40, 36), 5, "float32")])
def test_realign_as_frame(ctx, shape, n_res, dtype):
    """realignAsFrame (FrameWarpAligner.cpp:270-285) with the displacements add_frames found,
    on the raw frame as given, against the oracle's warps at the same D: the interpolated warp
    (float) and warpImageNormalised converted to CV_8U, CV_16U and CV_32F."""
    nD = n_res * shape[0] + 1
    st = synth.make_stack(shape, 2, nD, seed=21, n_blobs=120, sigma=(1.5, 4.0), dtype=dtype)
    al = F.Aligner(ctx, st["reference"], n_resampling_points=n_res, smoothing=1.0)
    D_g, _, _ = al.add_frames(list(st["frames"]), rel_delta=1e-4, max_iter=50, skip_repeat=1)
    orc = O.Oracle().set_reference(synth.with_offset(np.asarray(st["reference"], np.float32)), nD)
    for i, raw in enumerate(st["frames"]):
        got = al.realign_as_frame(D_g[i], raw, interpolate_missing=True)
        exp = FE.realign_as_frame(orc, raw, D_g[i], True)
        assert got.dtype == np.float32
        assert np.max(np.abs(got - exp)) <= 1e-5 * np.max(np.abs(exp))
        w32 = al.realign_as_frame(D_g[i], raw, cv_type=5)
        e32 = FE.realign_as_frame(orc, raw, D_g[i], False, 5)
        assert np.max(np.abs(w32 - e32)) <= 1e-5 * np.max(np.abs(e32))
        for cv_type, dt in ((0, np.uint8), (2, np.uint16)):
            got = al.realign_as_frame(D_g[i], raw, cv_type=cv_type)
            assert got.dtype == dt
            # the device's own float splat converted exactly; the oracle's within one count
            conv = FE.to_u8(w32) if cv_type == 0 else FE.to_u16(w32)
            assert np.array_equal(got, conv)
            diff = np.abs(got.astype(np.int64) - FE.realign_as_frame(orc, raw, D_g[i], False, cv_type).astype(np.int64))
            assert diff.max() <= 1
    with pytest.raises(F.FaError):
        al.realign_as_frame(D_g[0], st["frames"][0], cv_type=3)
    al.close()


def test_clear_temp_keeps_frames(ctx):
    """clearTemp (GpuFrameWarper.cpp:154-158) frees the cached working spaces; the registered
    frames stay usable and give the same results."""
    st = synth.make_stack((1, 64, 72), 2, 9, seed=22)
    ref = F.Reference(ctx, synth.with_offset(st["reference"]), 9)
    frs = [ref.register(synth.with_offset(f)) for f in st["frames"]]
    before = [f.align(rel_delta=1e-4, max_iter=50) for f in frs]
    D = np.random.default_rng(2).normal(0, 0.5, (9, 3))
    e0 = frs[0].error(D)
    ref.clear_temp()
    assert frs[0].error(D) == e0
    after = [f.align(rel_delta=1e-4, max_iter=50) for f in frs]
    for (x1, s1), (x2, s2) in zip(before, after):
        assert np.array_equal(x1, x2) and s1["evaluations"] == s2["evaluations"]
    ref.clear_temp()
    for f in frs:
        f.release()
    ref.clear_temp()  # no frames left
    ref.release()


def test_shift_pixel(ctx):
    shape, n_res = (3, 40, 36), 4
    nD = n_res * shape[0] + 1
    st = synth.make_stack(shape, 1, nD, seed=9, n_blobs=60, sigma=(1.5, 4.0))
    al = F.Aligner(ctx, st["reference"], n_resampling_points=n_res)
    orc = O.Oracle().set_reference(synth.with_offset(st["reference"]), nD)
    D = np.random.default_rng(3).normal(0, 2, (nD, 3))
    for (x, y, z) in [(3.7, 5.2, 0.0), (20.0, 30.9, 2.0), (35.5, 39.0, 1.0), (-1.0, 4.0, 0.0), (10.0, 10.0, 5.0)]:
        loc = orc.warp_point(D, int(x), int(y), int(z))
        assert np.allclose(al.shift_pixel(D, x, y, z), (x - loc[0], y - loc[1], z - loc[2]), rtol=0, atol=1e-12)
    al.close()
