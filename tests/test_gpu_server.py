"""Continuous batching: a frame's result is the same alone and when it joins a server's slots
while other frames (other references, 2-D and 3-D, uint16 and float) are in flight."""
import numpy as np
import pytest

import oracle as O
import fa_amd as F
from fa_amd import serve
from helpers import small_stack

pytestmark = pytest.mark.gpu

D_TOL = 0.02
KW = dict(rel_delta=1e-4, max_iter=50, skip_repeat=1)


@pytest.fixture(scope="module")
def ctx():
    c = F.Context(0)
    yield c
    c.close()


def _setup(ctx):
    a = small_stack((1, 96, 80), n_frames=5, nD=9, seed=41)
    b = small_stack((1, 64, 72), n_frames=4, nD=13, seed=42, dtype="uint16")
    c = small_stack((6, 40, 36), n_frames=3, nD=19, seed=43, n_blobs=120, sigma=(1.5, 4.0))
    ra = F.Reference(ctx, a["ref_w"], 9)
    rb = F.Reference(ctx, b["ref_w"], 13, interline=1.3)
    rc = F.Reference(ctx, c["ref_w"], 19)
    frames = ([ra.register(f) for f in a["frames_w"]] + [rb.register(f) for f in b["frames"]] +
              [rc.register(f) for f in c["frames_w"]])
    return (a, b, c), (ra, rb, rc), frames


def test_server_update_widths_in_one_launch(ctx):
    """References whose updates differ in width and subproblem form share the slot batches:
    9 nodes (eigendecomposition, one wave), 60 nodes (cyclic reduction, one wave) and 150 nodes
    (cyclic reduction, 256 threads), whose updates then run as one launch. Every result is the
    frame's alone."""
    from helpers import MANY_NODE_CASES
    name, shape, nD, scan, gen = MANY_NODE_CASES[0]
    a = small_stack((1, 96, 80), n_frames=3, nD=9, seed=45)
    d = small_stack((1, 96, 80), n_frames=3, nD=60, seed=46, walk_sigma=0.1, sin_amp=1.5)
    e = small_stack(shape, n_frames=2, nD=nD, seed=47, **gen)
    refs = [F.Reference(ctx, a["ref_w"], 9), F.Reference(ctx, d["ref_w"], 60),
            F.Reference(ctx, e["ref_w"], nD, **scan)]
    frames = ([refs[0].register(f) for f in a["frames_w"]] + [refs[1].register(f) for f in d["frames_w"]] +
              [refs[2].register(f) for f in e["frames_w"]])
    alone = [f.align(**KW) for f in frames]
    srv = F.Server(refs, slots=16, **KW)
    got = {}
    for i in (6, 0, 3, 7, 1, 4, 2, 5):
        srv.submit(frames[i], tag=i)
    steps = 0
    while len(got) < len(frames):
        _, _, res = srv.step(rounds=1 + steps % 3)
        steps += 1
        for tag, x, st in res:
            got[tag] = (x, st)
        assert steps < 10000
    for i, (x1, s1) in enumerate(alone):
        x2, s2 = got[i]
        assert np.array_equal(x1, x2), i
        assert (s1["evaluations"], s1["status"], s1["f_final"]) == (s2["evaluations"], s2["status"], s2["f_final"]), i
    srv.close()
    for f in frames:
        f.release()


@pytest.mark.parametrize("slots", [4, 16])
def test_server_same_alone_and_in_mix(ctx, slots):
    """4 slots: fewer than the frames, some wait in the queue; 16 slots: every frame admitted
    at once, the 2-D and 3-D frames in their own lanes (slot batches on their own streams)."""
    stacks, refs, frames = _setup(ctx)
    alone = [f.align(**KW) for f in frames]
    srv = F.Server(refs, slots=slots, **KW)
    order = [11, 0, 5, 9, 1, 6, 10, 2, 7, 3, 8, 4]
    got = {}
    k = 0
    steps = 0
    per_step = 2 if slots < 16 else len(order)  # 16 slots: all at once, over both slot batches
    while len(got) < len(frames):
        # frames arrive per_step at a time while others are in flight
        for _ in range(per_step):
            if k < len(order):
                i = order[k]
                srv.submit(frames[i], tag=i)
                k += 1
        _, _, res = srv.step(rounds=1 + steps % 3)
        steps += 1
        for tag, x, st in res:
            got[tag] = (x, st)
        assert steps < 10000
    for i, (x1, s1) in enumerate(alone):
        x2, s2 = got[i]
        assert np.array_equal(x1, x2), i
        assert (s1["evaluations"], s1["iterations"], s1["status"]) == \
            (s2["evaluations"], s2["iterations"], s2["status"]), i
        assert s1["f_final"] == s2["f_final"]
    # the 2-D float frames against the oracle
    orc = O.Oracle().set_reference(stacks[0]["ref_w"], 9)
    for i in range(2):
        x_o, _ = orc.align(stacks[0]["frames_w"][i], **KW)
        assert np.max(np.abs(got[i][0] - x_o)) <= D_TOL
    # a started run is independent of later submissions: run everything again, one at a time
    for i in (3, 10):
        srv.submit(frames[i], tag=100 + i)
        res = []
        while not res:
            _, _, res = srv.step(rounds=2)
        assert np.array_equal(res[0][1], alone[i][0])
    srv.close()


def test_server_lanes_same_alone(ctx):
    """Three lanes: 2-D frames, small 3-D volumes, and 3-D volumes with more than 16 times their
    voxels, each lane on its own stream and cadence. Frames arrive while the others are in
    flight; every result is the frame's alone."""
    a = small_stack((1, 64, 72), n_frames=4, nD=13, seed=61, dtype="uint16")
    c = small_stack((6, 40, 36), n_frames=3, nD=19, seed=62, n_blobs=120, sigma=(1.5, 4.0))
    d = small_stack((8, 128, 160), n_frames=2, nD=25, seed=63, n_blobs=600, sigma=(1.5, 4.0))
    refs = [F.Reference(ctx, a["ref_w"], 13), F.Reference(ctx, c["ref_w"], 19), F.Reference(ctx, d["ref_w"], 25)]
    frames = ([refs[0].register(f) for f in a["frames"]] + [refs[1].register(f) for f in c["frames_w"]] +
              [refs[2].register(f) for f in d["frames_w"]])
    alone = [f.align(**KW) for f in frames]
    srv = F.Server(refs, slots=8, **KW)
    order = [7, 0, 4, 8, 1, 5, 2, 6, 3]
    got = {}
    k = steps = 0
    while len(got) < len(frames):
        for _ in range(2):
            if k < len(order):
                srv.submit(frames[order[k]], tag=order[k])
                k += 1
        _, _, res = srv.step(rounds=2)
        steps += 1
        for tag, x, st in res:
            got[tag] = (x, st)
        assert steps < 20000
    for i, (x1, s1) in enumerate(alone):
        x2, s2 = got[i]
        assert np.array_equal(x1, x2), i
        assert (s1["evaluations"], s1["status"], s1["f_final"]) == (s2["evaluations"], s2["status"], s2["f_final"]), i
    srv.close()
    for f in frames:
        f.release()


def test_server_long_chain_same_alone(ctx):
    """A 3-D reference of 513 nodes (1539 parameters: its subproblems run in the batched
    eigen-space kernels) shares a lane with a 19-node one; frames arrive while others are in
    flight, and every result is the frame's alone."""
    from helpers import LONG_CHAIN_CASE
    _, shape, nD, scan, gen = LONG_CHAIN_CASE
    c = small_stack((6, 40, 36), n_frames=2, nD=19, seed=71, n_blobs=120, sigma=(1.5, 4.0))
    d = small_stack(shape, n_frames=3, nD=nD, seed=72, **gen)
    refs = [F.Reference(ctx, c["ref_w"], 19), F.Reference(ctx, d["ref_w"], nD, **scan)]
    frames = [refs[0].register(f) for f in c["frames_w"]] + [refs[1].register(f) for f in d["frames_w"]]
    alone = [f.align(**KW) for f in frames]
    srv = F.Server(refs, slots=4, **KW)
    order = [2, 0, 3, 1, 4]
    got = {}
    k = steps = 0
    while len(got) < len(frames):
        if k < len(order):
            srv.submit(frames[order[k]], tag=order[k])
            k += 1
        _, _, res = srv.step(rounds=2)
        steps += 1
        for tag, x, st in res:
            got[tag] = (x, st)
        assert steps < 20000
    for i, (x1, s1) in enumerate(alone):
        x2, s2 = got[i]
        assert np.array_equal(x1, x2), i
        assert (s1["evaluations"], s1["status"], s1["f_final"]) == (s2["evaluations"], s2["status"], s2["f_final"]), i
    srv.close()
    for f in frames:
        f.release()


def test_run_stream_error_drains_the_server(ctx):
    """A frame the server cannot take (its reference is not served) ends the native stream with
    the error after the frames submitted before it have run; the server is then empty and
    serves a following stream whose results equal the frames aligned alone."""
    st = small_stack((1, 64, 64), n_frames=3, nD=9, seed=51)
    other = small_stack((1, 64, 64), n_frames=1, nD=9, seed=52)
    r1 = F.Reference(ctx, st["ref_w"], 9)
    r2 = F.Reference(ctx, other["ref_w"], 9)
    srv = F.Server([r1], slots=4, **KW)
    frames = [np.ascontiguousarray(f) for f in st["frames_w"]]
    with pytest.raises(F.FaError):
        srv.run_stream([0.0, 0.0, 0.001], [r1, r1, r2], [frames[0], frames[1], other["frames_w"][0]])
    assert srv.step(1)[:2] == (0, 0)
    lat, res, cnt = srv.run_stream([0.0, 0.0005, 0.001], [r1, r1, r1], frames)
    assert np.all(lat >= 0) and cnt["steps"] > 0
    for f, (x, s) in zip(frames, res):
        fr = r1.register(f)
        xa, sa = fr.align(**KW)
        fr.release()
        assert np.array_equal(x, xa) and s["evaluations"] == sa["evaluations"]
    srv.close()
    r1._release_handle()
    r2._release_handle()


def test_run_stream_refuses_a_server_holding_frames(ctx):
    """The native stream matches results to its arrivals by tag, so a server that still holds
    a frame submitted before the call is refused (FA_E_ARG) and keeps that frame, whose result
    comes back through step as usual; the emptied server then takes the stream, and both
    results equal the frames aligned alone."""
    st = small_stack((1, 64, 64), n_frames=2, nD=9, seed=53)
    r1 = F.Reference(ctx, st["ref_w"], 9)
    srv = F.Server([r1], slots=4, **KW)
    frames = [np.ascontiguousarray(f) for f in st["frames_w"]]
    srv.submit_host(r1, frames[0], tag=7)
    with pytest.raises(F.FaError) as e:
        srv.run_stream([0.0], [r1], [frames[1]])
    assert e.value.code == -1
    got = []
    for _ in range(2000):
        a, q, out = srv.step(1)
        got += out
        if a == 0 and q == 0:
            break
    assert [t for t, _, _ in got] == [7]
    lat, res, cnt = srv.run_stream([0.0], [r1], [frames[1]])
    assert lat[0] >= 0 and cnt["steps"] > 0
    for f, x in zip(frames, [got[0][1], res[0][0]]):
        fr = r1.register(f)
        xa, _ = fr.align(**KW)
        fr.release()
        assert np.array_equal(x, xa)
    srv.close()
    r1._release_handle()


def test_server_errors(ctx):
    st = small_stack((1, 64, 64), n_frames=1, nD=9, seed=44)
    r1 = F.Reference(ctx, st["ref_w"], 9)
    r2 = F.Reference(ctx, st["ref_w"], 9)
    srv = F.Server([r1], slots=2)
    fr = r2.register(st["frames_w"][0])
    with pytest.raises(F.FaError):
        srv.submit(fr, tag=1)
    fr.release()
    with pytest.raises(F.FaError):
        r1._release_handle()  # the server still serves it
    srv.close()
    r1._release_handle()
    r2._release_handle()


@pytest.mark.parametrize("slots", [6, 16])
def test_server_host_frames_and_arena(ctx, slots):
    """Host frames uploaded into the arena give the same results as registered frames; an
    arena that holds two frames at a time still serves a queue of many (one slot batch); with
    16 slots every frame is admitted at once, spread over both slot batches."""
    stacks, refs, frames = _setup(ctx)
    alone = [f.align(**KW) for f in frames]
    hosts = ([(refs[0], f) for f in stacks[0]["frames_w"]] + [(refs[1], f) for f in stacks[1]["frames"]] +
             [(refs[2], f) for f in stacks[2]["frames_w"]])
    big = max(f.nbytes for _, f in hosts)
    # 6 slots: an arena of two frames; 16 slots: all frames admitted at once, over both parts
    srv = F.Server(refs, slots=slots, arena_bytes=(2 if slots < 16 else len(hosts)) * (big + 512), **KW)
    pinned = [F.PinnedArray(f.shape, f.dtype) for _, f in hosts]
    for p, (_, f) in zip(pinned, hosts):
        p.array[...] = f
    for i, ((r, _), p) in enumerate(zip(hosts, pinned)):
        # float frames are already offset, uint16 counts get +1 on the device
        srv.submit_host(r, p.array, tag=i, offset=0.0 if p.array.dtype == np.float32 else None)
    got = {}
    steps = 0
    while len(got) < len(hosts):
        _, _, res = srv.step(rounds=2)
        used, total = srv.memory()
        assert used <= total
        for tag, x, st in res:
            got[tag] = x
        steps += 1
        assert steps < 10000
    for i, (x1, _) in enumerate(alone):
        assert np.array_equal(x1, got[i]), i
    assert srv.memory()[0] == 0
    with pytest.raises(F.FaError):  # never fits
        F.Server(refs[:1], slots=1, arena_bytes=256).submit_host(refs[0], pinned[0].array, tag=0)
    srv.close()
    for p in pinned:
        p.close()


@pytest.mark.parametrize("native", [False, True])
def test_serve_loop_small(ctx, native):
    """The request loop here and in the library (fa_server_run_stream): every served result is
    the frame aligned alone."""
    acqs = serve.make_acquisitions(seed=5, shapes=[(1, 64, 64), (1, 96, 80), (4, 40, 36)], pool=2,
                                   node_range=(8, 12))
    refs = [F.Reference(ctx, serve.synth.with_offset(a["reference"]), a["nD"]) for a in acqs]
    pools = [a["frames"] for a in acqs]
    srv = F.Server(refs, slots=8, arena_bytes=4 * max(p[0].nbytes for p in pools), **KW)
    times, acq = serve.arrivals(len(acqs), rate=400.0, duration=0.1, seed=6)
    loop = serve.ServeLoop(srv, refs, pools, rounds=2)
    keep = set(range(len(times)))
    res = loop.run(times, acq, keep=keep, native=native)
    assert np.all(np.isfinite(res["latency"])) and np.all(res["latency"] >= 0)
    # each served result is the same as aligning that frame alone
    for i in range(len(times)):
        a = int(acq[i])
        fr = refs[a].register(pools[a][i % len(pools[a])])
        x, st = fr.align(**KW)
        assert np.array_equal(x, res["kept"][i][0]), i
        fr.release()
    s = serve.summarize(res, times, acq, pools)
    assert s["frames"] == len(times)
    loop.close()
    srv.close()
