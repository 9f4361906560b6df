"""Benchmark: aligned frames/s on the 512x512, 30-node workload on one MI355X.

Workload: 2-D 512x512 uint16 photon-count frames (600-blob
phantom, mean ~50 photons/px, Poisson noise), 200 frames, 30 resampling points
(nD = 31 nodes, 2 components), trust region to 1e-4 relative change of the SSD
or 50 iterations. A step aligns the whole 200-frame batch from the zero start,
with frames already resident in HBM.

Prints one JSON line: value = frames/s of the batched form; roofline of the fused
warp+SSD+Jtr pass (algorithmic bytes / HIP-event-timed kernel time); the plain
one-frame-per-launch path and per-frame latency beside it; the CPU baseline (the
oracle built with the reference's Release flags, dense subproblem solves, one frame
per thread) timed on the same frames on this host, and the 256x256, 20-frame
workload on both; the checker build of the oracle gives the parity check of the
measured results (<= 0.02 px).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "frame-aligner_amd"))

import numpy as np  # noqa: E402

METRIC = "aligned frames/s @512²,30 nodes (≤0.02px err); p50/p99 latency; HBM GB/s"
HBM_PEAK_GBS = 8000.0
PROFILES = os.path.join(ROOT, "profiles")


def profile_facts(name):
    """`key: value` lines of a committed rocprof summary under profiles/ (HBM traffic per
    launch from FETCH_SIZE/WRITE_SIZE scaled by a copy of known size, the commit the counters
    were taken at, the measured limiter), or {}."""
    out = {}
    try:
        for line in open(os.path.join(PROFILES, name)):
            k, sep, v = line.partition(":")
            if sep and k.strip() in ("traffic_bytes_per_launch", "algorithmic_bytes_per_launch",
                                      "commit", "limiter", "workload"):
                v = v.strip()
                try:
                    v = float(v)
                except ValueError:
                    pass
                out[k.strip()] = v
    except OSError:
        pass
    return out


def traffic_of(name, achieved_gbs):
    """The roofline's `traffic` (HBM bytes per launch, measured by rocprof for the same
    workload) and its provenance, or (None, None)."""
    f = profile_facts(name)
    t = f.get("traffic_bytes_per_launch")
    if t is None:
        return None, None
    src = {"file": "profiles/" + name, "commit": f.get("commit"), "workload": f.get("workload")}
    a = f.get("algorithmic_bytes_per_launch")
    if a:
        src["dram_over_algorithmic"] = round(t / a, 3)
        # DRAM bytes moved at the measured launch rate, against the 8 TB/s peak
        src["dram_frac_of_peak"] = round(achieved_gbs * t / a / HBM_PEAK_GBS, 4)
    if f.get("limiter"):
        src["limiter"] = f["limiter"]
    return t, src


def parse_tuning(text):
    """`key=value,...` over the launch-shape fields (lds_floats, blocks_per_cu, use_lds,
    check_every, use_graph, tr_lds)."""
    return {k: int(v) for k, v in (kv.split("=") for kv in text.split(",") if kv)}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


# a single-GPU job's share of the host's cores: the CPU baseline uses at most this many threads
HOST_THREAD_SHARE = 16


def host_cpu(requested=0):
    """The host's CPU model and hardware_concurrency, and the worker threads the CPU baseline
    uses: hardware_concurrency (the reference's MEX pool, FrameAlignerMex.cpp:24-25), limited
    to the cores this process may run on and to a single-GPU job's share of the host."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    hc = os.cpu_count() or 1
    try:
        allowed = len(os.sched_getaffinity(0))
    except AttributeError:
        allowed = hc
    threads = requested or min(hc, allowed, HOST_THREAD_SHARE)
    return {"cpu_model": model, "hardware_concurrency": hc, "affinity_cores": allowed, "threads": threads,
            "thread_rule": f"min(hardware_concurrency, affinity, {HOST_THREAD_SHARE}: one GPU's share of the host)"
            if not requested else "--cpu-threads"}


# CPU baseline build and solve: the oracle compiled with the reference's Release flags,
# factorising the dense n x n B at every subproblem iteration as dlib does
CPU_BASELINE_BUILD = "oracle/build/libfa_oracle_fast.so: gcc -O3 -ffast-math -ftree-vectorize -msse3 (CMakeLists.txt:18-19)"
CPU_BASELINE_SOLVE = "dense Cholesky of the n x n B per subproblem iteration (dlib general_matrix, OptimisationModel.cpp:55-63)"


def cpu_time(O, ref_w, nD, frames, threads, tr):
    """The Release-flag oracle (dense subproblem solves) on frames, one frame per thread."""
    orc = O.Oracle(fast=True).set_reference(ref_w, nD)
    t = time.perf_counter()
    x, s, secs = orc.align_many(frames, threads=threads, dense=1, **tr)
    return x, s, secs, time.perf_counter() - t


def copy_probe(mib):
    """One plain device-to-device copy of mib MiB (hipMemcpy), whose rocprof FETCH_SIZE and
    WRITE_SIZE give the scale of those counters for a known byte count."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    n = mib << 20
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(n)) == 0 and hip.hipMalloc(C.byref(b), C.c_size_t(n)) == 0
    hip.hipMemset(a, 1, C.c_size_t(n))
    hip.hipDeviceSynchronize()
    assert hip.hipMemcpy(b, a, C.c_size_t(n), 3) == 0  # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    return mib


def isolated_pass(F, frames, tr, reps=10):
    """Mean duration of one fused pass over all frames (a one-part batch, every frame active,
    max_evals = 1 so each run is exactly one pass), from HIP events on the batch's stream."""
    b = F.Batch(frames[:63], **dict(tr, max_evals=1))  # < 64 frames: one stream, no overlap
    b.set_tuning(check_every=1)  # one round per poll: every timed launch is the full pass
    b.set_timing(True)
    b.run()
    ms, nbytes, k = 0.0, 0.0, 0
    for _ in range(reps):
        _, _, tm = b.run()
        ms += tm["pass_ms"]
        nbytes += tm["pass_bytes"]
        k += tm["pass_launches"]
    b.close()
    return ms / k, nbytes / k


def serving(F, args, ctx):
    """Serving form: 8 acquisitions (256^2 x 1 .. 1024^2 x 8, 10-40 nodes, uint16 blob
    phantoms) streaming frames with exponential inter-arrival times at args.serve_rate
    frames/s for args.serve_seconds, aligned by the continuous-batching server. Latency is
    arrival -> result. A sample of served results is checked against the same frame aligned
    alone (bit-identical) and, for the 256^2 2-D acquisition, against the CPU oracle."""
    from fa_amd import serve
    native = args.serve_loop == "native"
    t0 = time.time()
    acqs = serve.make_acquisitions(seed=args.seed, pool=2)
    log(f"serving: generated {len(acqs)} acquisitions in {time.time() - t0:.1f}s")
    refs = [F.Reference(ctx, serve.synth.with_offset(a["reference"]), a["nD"]) for a in acqs]
    pools = [a["frames"] for a in acqs]
    tr = dict(rel_delta=1e-4, max_iter=50, skip_repeat=1)
    srv = F.Server(refs, slots=args.serve_slots, **tr)
    loop = serve.ServeLoop(srv, refs, pools, rounds=args.serve_rounds)
    times, acq = serve.arrivals(len(acqs), args.serve_rate, args.serve_seconds, seed=args.seed + 1)
    loop.run(times[:200], acq[:200], native=native)  # warm-up
    keep = set(range(0, len(times), max(1, len(times) // 64)))
    res = loop.run(times, acq, keep=keep, native=native)
    out = serve.summarize(res, times, acq, pools)
    out["workload"] = ("8 acquisitions " + ", ".join("x".join(map(str, a["shape"][::-1])) + f"/{a['nD']}n"
                                                     for a in acqs) + f", Poisson arrivals {args.serve_rate:g}/s")
    out["slots"] = args.serve_slots
    out["rounds_per_step"] = args.serve_rounds
    out["loop"] = args.serve_loop
    # the same stream at lower offered rates: the highest the GPU sustains is where the
    # latency stops growing with the run's length
    out["at_lower_rates"] = []
    for rate in [float(r) for r in args.serve_lower.split(",") if r]:
        t2, a2 = serve.arrivals(len(acqs), rate, args.serve_seconds, seed=args.seed + 2)
        r2 = loop.run(t2, a2, native=native)
        s2 = serve.summarize(r2, t2, a2, pools)
        out["at_lower_rates"].append({"offered_frames_per_s": s2["offered_frames_per_s"],
                                      "frames_per_s": s2["frames_per_s"], "latency_ms": s2["latency_ms"],
                                      "max_queued": s2["max_queued"]})
        log(f"serving at {rate:g}/s: {s2['frames_per_s']} frames/s, p50 {s2['latency_ms']['p50']} ms, "
            f"p99 {s2['latency_ms']['p99']} ms")
    # served == alone, on a sample
    same = True
    for i in sorted(keep):
        a = int(acq[i])
        fr = refs[a].register(pools[a][i % len(pools[a])])
        x, _ = fr.align(**tr)
        same = same and bool(np.array_equal(x, res["kept"][i][0]))
        fr.release()
    out["same_as_alone"] = {"frames": len(keep), "all_identical": same}
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        a0 = 0  # the 256^2 2-D acquisition
        orc = O.Oracle().set_reference(serve.synth.with_offset(acqs[a0]["reference"]), acqs[a0]["nD"])
        err = 0.0
        n = 0
        for i in sorted(keep):
            if int(acq[i]) != a0:
                continue
            x_o, _ = orc.align(serve.synth.with_offset(pools[a0][i % len(pools[a0])]), **tr)
            err = max(err, float(np.max(np.abs(res["kept"][i][0] - x_o))))
            n += 1
        out["max_disp_err_px_vs_oracle"] = {"frames": n, "acquisition": "256x256x1", "max": err}
    loop.close()
    srv.close()
    for r in refs:
        r.release()
    return out


def serving_512(F, args, ctx, plain_fps):
    """The serving form on the headline workload: one acquisition of 512x512 uint16 frames,
    nD = 31, streamed with exponential inter-arrival times through the continuous-batching
    server at 3x the plain path's frames/s (the batched-serving target's operating point) and at
    args.serve512_high; frames/s and p50/p99 arrival -> result latency, and served == alone on a
    sample."""
    native = args.serve_loop == "native"
    from fa_amd import serve, synth
    st = synth.make_stack((1, 512, 512), 64, 31, seed=args.seed + 5, n_blobs=600, mean=50.0, dtype="uint16")
    ref = F.Reference(ctx, synth.with_offset(st["reference"]), 31)
    tr = dict(rel_delta=1e-4, max_iter=50, skip_repeat=1)
    srv = F.Server([ref], slots=args.serve_slots, **tr)
    loop = serve.ServeLoop(srv, [ref], [st["frames"]], rounds=args.serve_rounds)
    out = {"workload": "512x512 uint16, nD=31, one acquisition, 64 distinct frames cycled, Poisson arrivals",
           "plain_frames_per_s": round(plain_fps, 1), "loop": args.serve_loop, "runs": []}
    w, wa = serve.arrivals(1, 2000.0, 0.2, seed=args.seed + 6)
    loop.run(w, wa, native=native)  # warm-up
    for rate, dur in ((3.0 * plain_fps, args.serve_seconds), (args.serve512_high, args.serve_seconds),
                      (args.serve512_sat, min(args.serve_seconds, 0.5))):
        if not rate:
            continue
        t, a = serve.arrivals(1, rate, dur, seed=args.seed + 7)
        keep = set(range(0, len(t), max(1, len(t) // 16)))
        r = loop.run(t, a, keep=keep, native=native)
        s = serve.summarize(r, t, a, [st["frames"]])
        same = True
        for i in sorted(keep):
            fr = ref.register(st["frames"][i % 64])
            x, _ = fr.align(**tr)
            same = same and bool(np.array_equal(x, r["kept"][i][0]))
            fr.release()
        # every frame is uploaded from pinned host memory when admitted
        out["runs"].append({"offered_frames_per_s": s["offered_frames_per_s"], "frames_per_s": s["frames_per_s"],
                            "vs_plain": round(s["frames_per_s"] / plain_fps, 2), "latency_ms": s["latency_ms"],
                            "max_in_flight": s["max_in_flight"], "max_queued": s["max_queued"],
                            "upload_gb_per_s": round(s["frames_per_s"] * st["frames"][0].nbytes / 1e9, 2),
                            "same_as_alone": same})
        log(f"serving 512^2 at {rate:.0f}/s: {s['frames_per_s']} frames/s, p50 {s['latency_ms']['p50']} ms, "
            f"p99 {s['latency_ms']['p99']} ms")
    loop.close()
    srv.close()
    ref.release()
    return out


def timed_steps(batch, steps):
    """`steps` untimed runs (graph replay where tuned), then `steps` runs with HIP events around
    every fused pass. Returns (results, stats, wall s per run, pass timing sums)."""
    t = time.perf_counter()
    for _ in range(steps):
        xs, stats, _ = batch.run()
    wall = (time.perf_counter() - t) / steps
    batch.set_timing(True)
    tot = {"pass_ms": 0.0, "pass_bytes": 0.0, "pass_px": 0.0, "pass_launches": 0}
    for _ in range(steps):
        _, _, tm = batch.run()
        for k in tot:
            tot[k] += tm[k]
    batch.set_timing(False)
    return xs, stats, wall, tot


# bytes per pixel the pass reads beyond the algorithmic ones: the node fraction (3-D, its own
# array) or the 2-D record's fourth float (the node fraction in the padding)
RECORD_EXTRA_BYTES_PER_PX = 4


def roofline_of(tm, kernel, traffic_file=None):
    """Pass roofline from HIP events around every fused pass (eager runs; a batch split into
    parts runs them one after the other in those runs, so each pass is timed alone). achieved =
    algorithmic bytes (per active pixel: frame sample, reference intensity, n_dim gradient
    components) / summed pass time; the bytes the pass actually requests (plus the node
    fraction) are reported beside it."""
    ms, launches = tm["pass_ms"], max(int(tm["pass_launches"]), 1)
    gbs = tm["pass_bytes"] / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
    rec = (tm["pass_bytes"] + RECORD_EXTRA_BYTES_PER_PX * tm["pass_px"]) / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
    out = {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": None, "kernel": kernel,
           "bytes_per_px": "s_f + 4 + 4 n_dim (frame, reference, gradient)",
           "launches": int(tm["pass_launches"]), "avg_launch_us": round(1e3 * ms / launches, 2),
           "algorithmic_bytes_per_launch": round(tm["pass_bytes"] / launches),
           "requested_bytes": {"achieved": round(rec, 1), "frac": round(rec / HBM_PEAK_GBS, 4),
                               "bytes_per_px": "algorithmic + 4 (node fraction)"}}
    if traffic_file:
        out["traffic"], src = traffic_of(traffic_file, gbs)
        if src:
            out["traffic_source"] = src
    return out


def precision_3d(F, args, ctx):
    """Precision form (3-D): 256x256x32 float32 volumes (1500-blob phantom, sigma 1.5-4 voxels,
    Poisson noise), 16 resampling points per plane (nD = 16*32+1 = 513), 3 components (z within
    2 voxels); FP64 sample positions and per-lane sums in the pass, FP64 cross-block reductions
    and block-tridiagonal solves. Frames are resampled with trilinear interpolation (order 1) to
    bound the generator's time. Reported: volumes/s of the batched run, the 3-D pass roofline,
    and per result the largest displacement difference to the FP64 oracle and to the
    float-storage oracle on a sample of volumes; the CPU baseline times the Release-flag oracle
    build with dense solves on that sample."""
    from fa_amd import synth
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    cfg = synth.CONFIGS["3d_256x256x32_f32"]
    t0 = time.time()
    st = synth.make_stack(cfg["shape"], args.c3_frames, cfg["nD"], seed=args.seed + 3, n_blobs=cfg["n_blobs"],
                          sigma=cfg["sigma"], dtype="float32", order=1)
    log(f"precision_3d: generated {args.c3_frames} volumes in {time.time() - t0:.1f}s")
    ref_w = synth.with_offset(st["reference"])
    tr = dict(rel_delta=1e-4, max_iter=50, skip_repeat=1)
    ref = F.Reference(ctx, ref_w, cfg["nD"])
    frames = [ref.register(synth.with_offset(f)) for f in st["frames"]]
    b = F.Batch(frames, **tr)
    b.run()
    xs, stats, wall, tm = timed_steps(b, args.steps)
    n = len(frames)
    out = {"workload": f"256x256x32 float32, {n} volumes, nD=513 x 3 components, TR to 1e-4 rel or 50 "
                       f"iterations, batched", "volumes": n,
           "volumes_per_s": round(n / wall, 2), "ms_per_step": round(1e3 * wall, 3),
           "mvox_per_s": round(n * ref.n_px / wall / 1e6, 1),
           "evaluations_per_volume": {"mean": float(np.mean([s_["evaluations"] for s_ in stats])),
                                      "max": int(np.max([s_["evaluations"] for s_ in stats]))},
           "roofline": roofline_of(tm, "k_pass<3> (fused trilinear warp + SSD + Jtr)",
                                   "k_pass3_traffic_256x256x32_50v_nD513.txt")}
    # the pass with every volume active (the first round's pass), beside the average over the
    # run's launches, whose later rounds carry fewer volumes
    iso_ms, iso_bytes = isolated_pass(F, frames, tr, reps=4)
    out["roofline"]["all_active"] = {"avg_launch_us": round(1e3 * iso_ms, 2), "bytes_per_launch": iso_bytes,
                                     "achieved": round(iso_bytes / (iso_ms * 1e-3) / 1e9, 1),
                                     "frac": round(iso_bytes / (iso_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    if not args.no_cpu:
        k = min(args.c3_oracle, n)
        host = host_cpu(args.cpu_threads)
        threads = min(k, host["threads"])
        fw = np.stack([synth.with_offset(f) for f in st["frames"][:k]])
        per, xo, so, orcs = [], [], [], []
        for fp64 in (True, False):
            orc = O.Oracle(fp64=fp64).set_reference(ref_w, cfg["nD"])
            orcs.append(orc)
            x_o, s_o, _ = orc.align_many(fw, threads=threads, **tr)
            per.append([float(np.max(np.abs(xs[i] - x_o[i]))) for i in range(k)])
            xo.append(x_o)
            so.append(s_o)
        # baseline: the Release-flag build with the reference's dense solves (n = 1539)
        _, s_f, secs, wall_o = cpu_time(O, ref_w, cfg["nD"], fw, threads, tr)
        out["cpu_baseline"] = {"value": round(k / wall_o, 3), "unit": "volumes/s", "cores": threads,
                               "kind": "port", "host": host, "build": CPU_BASELINE_BUILD,
                               "solve": CPU_BASELINE_SOLVE,
                               "sample": f"first {k} of the {n} volumes, one volume per thread, "
                                         f"{wall_o:.1f}s wall, {secs.sum():.1f}s CPU",
                               "evaluations_per_volume": float(np.mean([d["evaluations"] for d in s_f]))}
        out["max_disp_err_px_per_result"] = {"vs_fp64_oracle": [round(e, 6) for e in per[0]],
                                             "vs_float_oracle": [round(e, 6) for e in per[1]]}
        # how far the oracle's own two builds end apart on the same volumes, and the final SSD
        # each run reaches (relative to the FP64 oracle's): where the builds disagree by pixels,
        # the objective does not fix D to 0.02 px
        out["float_oracle_vs_fp64_oracle_px"] = [round(float(np.max(np.abs(xo[1][i] - xo[0][i]))), 6)
                                                 for i in range(k)]
        out["final_ssd_rel_to_fp64_oracle"] = {
            "gpu": [round(stats[i]["f_final"] / so[0][i]["f_final"] - 1, 8) for i in range(k)],
            "float_oracle": [round(so[1][i]["f_final"] / so[0][i]["f_final"] - 1, 8) for i in range(k)]}
        # one evaluation at the generator's displacements: the GPU's SSD and Jtr (FP32 warp and
        # compensated sums, FP64 across work items) against the FP64 oracle's, and JtJ
        ssd_e, jtr_e = [], []
        for i in range(k):
            D = st["D_true"][i]
            e_g, j_g = frames[i].jacobian(D)
            e_o, j_o = orcs[0].get_jacobian(fw[i], D)
            ssd_e.append(abs(e_g - e_o) / abs(e_o))
            jtr_e.append(float(np.linalg.norm(j_g - j_o) / np.linalg.norm(j_o)))
        H_g, H_o = ref.H, orcs[0].hessian()
        out["at_true_D_vs_fp64_oracle"] = {
            "ssd_rel_err": [float(f"{e:.3g}") for e in ssd_e], "jtr_rel_err": [float(f"{e:.3g}") for e in jtr_e],
            "jtj_rel_err": float(f"{np.linalg.norm(H_g - H_o) / np.linalg.norm(H_o):.3g}"),
            "tolerance": {"ssd": 1e-4, "jtr_jtj": 1e-3}}
    b.close()
    for f in frames:
        f.release()
    ref.release()
    return out


def uneven_batch(F, args, ctx):
    """Batched form: 64 independent stacks of uneven size in one batch (hipGraph-replayed
    trust-region rounds, launch shape from the autotuner). Sizes uniform over {128^2, 256^2,
    512^2, 1024^2} px x {1, 4, 16} planes, 10-40 resampling points per plane (nD = n * planes + 1),
    20-100 frames per stack, float32 blob phantoms with a seeded per-stack blob density. Each
    frame is the stack's noise-free phantom pulled through its own node displacement (the
    generator's motion model), computed with the library's interpolated warp on the device."""
    from fa_amd import synth, tune
    rng = np.random.default_rng(args.seed + 4)
    t0 = time.time()
    stacks, refs, frames = [], [], []
    tr = dict(rel_delta=1e-4, max_iter=50, skip_repeat=1)
    for i in range(args.c4_stacks):
        side = int(rng.choice([128, 256, 512, 1024]))
        nz = int(rng.choice([1, 4, 16]))
        n_res = int(rng.integers(10, 41))
        nD = n_res * nz + 1
        n_frames = int(rng.integers(20, 101))
        shape = (nz, side, side)
        density = float(rng.uniform(300, 900))  # blobs per 512^2 px area per plane group
        n_blobs = max(20, int(density * side * side / 512 ** 2 * max(1.0, nz / 4)))
        srng = np.random.default_rng(rng.integers(1 << 62))
        ph = synth.blob_phantom(shape, n_blobs, (2.0, 6.0) if nz == 1 else (1.5, 4.0), rng=srng)
        ph32 = synth.with_offset(ph)
        ref = F.Reference(ctx, ph32, nD)
        src = ref.register(ph32)  # the phantom as a frame, to pull-sample it through -D
        bg = float(np.median(ph32))
        n_dim = 3 if nz > 1 else 2
        for _ in range(n_frames):
            D = synth.node_displacements(nD, n_dim, srng)
            frames.append(ref.register(src.warp_interpolated(-D, invalid_value=bg)))
        src.release()
        refs.append(ref)
        stacks.append(f"{side}x{side}x{nz}/{nD}n/{n_frames}f")
    n = len(frames)
    vox = sum(f.ref.n_px for f in frames)
    log(f"uneven_batch: {len(stacks)} stacks, {n} frames, {vox / 1e9:.2f} Gvoxel, generated in {time.time() - t0:.1f}s")
    b = F.Batch(frames, **tr)
    best = None
    if args.set_tuning:
        b.set_tuning(**parse_tuning(args.set_tuning))
    elif not args.no_tune:
        best, _, trials = tune.tune_batch(b, reps=2, sweeps=1)
        log(f"uneven_batch: autotuner {best} ({len(trials)} shapes)")
    b.run()
    xs, stats, wall, tm = timed_steps(b, max(1, args.steps - 1))
    out = {"workload": f"{len(stacks)} stacks, {n} frames, {vox / 1e9:.2f} Gvoxel per pass over all frames",
           "stacks": stacks, "frames": n, "frames_per_s": round(n / wall, 1), "ms_per_step": round(1e3 * wall, 2),
           "mvox_per_s": round(vox / wall / 1e6, 1), "tuning": b.tuning,
           "launch": "one batch whose trust-region rounds (fused pass per launch class, update) are replayed "
                     "as captured hipGraphs between host polls, not one persistent launch: a persistent "
                     "grid with a static split of the items left a tail (94 vs 67 us per 100-frame pass) and "
                     "an atomic work queue raised the pass's registers",
           "evaluations_per_frame": {"mean": float(np.mean([s_["evaluations"] for s_ in stats])),
                                     "max": int(np.max([s_["evaluations"] for s_ in stats]))},
           "roofline": roofline_of(tm, "k_pass<2>, k_pass<3>")}
    b.close()
    for f in frames:
        f.release()
    for r in refs:
        r.release()
    return out


def end_to_end(F, args, ctx, st, tr):
    """The headline workload through the aligner front end (FrameWarpAligner::addFrame,
    FrameWarpAligner.cpp:117-223) for all frames at once: raw uint16 frames uploaded from host
    memory, +1 (no blur), the three starts (zero twice and the rigid phase-correlation shift) with
    the best one taken, the trust region, and the realigned uint16 frames copied back. frames/s
    over whole calls, and the host-device bytes they move."""
    al = F.Aligner(ctx, st["reference"], type="warp", n_resampling_points=30)
    frames = list(st["frames"])
    al.add_frames(frames, outputs=("realigned",), **tr)  # plans, device arrays and pinned staging
    reps, t = max(1, args.steps), time.perf_counter()
    for _ in range(reps):
        D, res, imgs = al.add_frames(frames, outputs=("realigned",), **tr)
    wall = (time.perf_counter() - t) / reps
    al.close()
    n = len(frames)
    moved = n * frames[0].nbytes + imgs["realigned"].nbytes
    starts = [r["start"] for r in res]
    return {"workload": f"512x512 uint16, {n} raw frames per call, 30 resampling points (nD=31), three starts "
                        f"with the rigid FFT start, TR to 1e-4 rel or 50 iterations, realigned frames returned",
            "frames_per_s": round(n / wall, 1), "ms_per_call": round(1e3 * wall, 2), "calls": reps,
            "host_device_bytes_per_call": int(moved), "host_device_gb_per_s": round(moved / wall / 1e9, 2),
            "start_chosen": {str(k): starts.count(k) for k in sorted(set(starts))}}


def large_working_set(F, args, ref, raw, tuning, tr, xs_ref):
    """The headline workload with args.large_frames frames resident at once (the 200 distinct
    frames each registered several times, so every copy is its own device buffer): frames,
    reference record and work-item partials exceed the 256 MB last-level cache (MALL), so the
    pass streams its frames from HBM. frames/s, the pass roofline by HIP events, and whether
    every copy's result equals the 200-frame batch's for the same frame."""
    n = args.large_frames
    frames = [ref.register(raw[i % len(raw)]) for i in range(n)]
    b = F.Batch(frames, **tr)
    b.set_tuning(**tuning)
    b.run()
    xs, _, wall, tm = timed_steps(b, args.steps)
    same = all(np.array_equal(xs[i], xs_ref[i % len(raw)]) for i in range(n))
    b.close()
    for f in frames:
        f.release()
    frame_bytes = sum(f.nbytes for f in raw[:1]) * n
    return {"workload": f"512x512 uint16, {n} frames ({len(raw)} distinct, each a separate device copy), "
                        f"nD=31, TR to 1e-4 rel or 50 iterations, batched",
            "frames": n, "frame_bytes": int(frame_bytes), "frames_per_s": round(n / wall, 1),
            "ms_per_step": round(1e3 * wall, 3), "same_result_as_200_frame_batch": same,
            "roofline": roofline_of(tm, "k_pass<2> (fused bilinear warp + SSD + Jtr)",
                                    f"k_pass_traffic_512x512_{n}f_nD31.txt")}


def headline(F, args, ctx):
    """The 512x512 uint16, 200-frame, nD = 31 workload: batched frames/s (the `value`), the pass
    roofline, the plain one-frame-per-launch path and the CPU oracle beside it."""
    from fa_amd import synth, tune

    nD, shape = 31, (1, 512, 512)
    t0 = time.time()
    st = synth.make_stack(shape, args.frames, nD, seed=args.seed, n_blobs=600, mean=50.0,
                          dtype="uint16")
    log(f"generated {args.frames} frames in {time.time() - t0:.1f}s")
    ref_w = synth.with_offset(st["reference"])
    tr = dict(rel_delta=1e-4, max_iter=50, skip_repeat=1)

    chunk = args.chunk
    chunk_trials = None
    if not chunk and not args.no_tune:
        # work-item size (pixels per workgroup tile), then the launch shape, by the autotuner
        nt = args.tune_frames or args.frames
        chunk, chunk_trials = tune.tune_chunk(ctx, ref_w, nD, st["frames"][:nt], tr=tr, reps=5)
        log(f"autotuner: chunk_px {chunk} ({ {k: round(v * 1e3, 3) for k, v in chunk_trials.items()} } ms)")
    ref = F.Reference(ctx, ref_w, nD, chunk_px=chunk)
    frames = [ref.register(f) for f in st["frames"]]  # uint16 counts, +1 on the device
    batch = F.Batch(frames, **tr)
    if args.set_tuning:
        batch.set_tuning(**parse_tuning(args.set_tuning))
    elif not args.no_tune:
        best, best_t, trials = tune.tune_batch(batch, reps=5, sweeps=1)
        log(f"autotuner: {best} ({best_t * 1e3:.3f} ms per run, {len(trials)} shapes tried)")
    tuning = batch.tuning
    for _ in range(args.warmup):
        batch.run()
    # timed region of `value`: K untimed steps (hipGraph replay when the tuner chose it)
    tot_wall = 0.0
    for k in range(args.steps):
        t = time.perf_counter()
        xs, stats, _ = batch.run()
        tot_wall += time.perf_counter() - t
        log(f"step {k}: {1e3 * (time.perf_counter() - t):.2f} ms")
    # the same steps again with HIP events around every fused pass (eager launches, since
    # events inside a replayed graph cannot be read): per-pass kernel time and bytes
    batch.set_timing(True)
    tsum = {"pass_ms": 0.0, "pass_bytes": 0.0, "pass_px": 0.0, "pass_launches": 0}
    timed_wall = 0.0
    for k in range(args.steps):
        t = time.perf_counter()
        xs_t, _, tm = batch.run()
        timed_wall += time.perf_counter() - t
        for f in tsum:
            tsum[f] += tm[f]
        log(f"timed step {k}: {tm['wall_ms']:.1f} ms, {tm['pass_launches']} passes, "
            f"pass {tm['pass_ms']:.2f} ms")
    batch.set_timing(False)
    n = args.frames
    fps = n * args.steps / tot_wall
    ms_per_step = 1e3 * tot_wall / args.steps
    evals = [s["evaluations"] for s in stats]
    same_timed = all(np.array_equal(a, b) for a, b in zip(xs, xs_t))

    # plain path: one frame per launch sequence
    lat = []
    npl = max(1, min(args.plain_frames, n))
    for i in range(npl):  # creates each frame's one-frame batch outside the timed loop
        frames[i].align(**tr)
    x_plain = []
    for i in range(npl):
        t = time.perf_counter()
        x, _ = frames[i].align(**tr)
        lat.append(time.perf_counter() - t)
        x_plain.append(x)
    lat = np.array(lat) * 1e3
    plain_fps = npl / (lat.sum() * 1e-3)
    plain_same = all(np.array_equal(a, b) for a, b in zip(x_plain, xs[:npl]))

    # the pass alone, all frames active, no other work on the GPU: one fused pass over the
    # whole batch per launch, timed with HIP events
    iso_ms, iso_bytes = isolated_pass(F, frames, tr) if not args.no_isolated else (0.0, 0.0)

    out = {
        "metric": METRIC,
        "value": round(fps, 2),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded blob phantom, Poisson counts; no dataset)",
        "config": {"workload": "512x512 uint16, 200 frames, 30 resampling points (nD=31), "
                               "TR to 1e-4 rel or 50 iterations, batched",
                   "frames": n, "nD": nD, "shape": list(shape), "chunk_px": chunk,
                   "tuning": tuning, "tuned": not args.no_tune and not args.set_tuning},
        "timed_ms_per_step": round(1e3 * timed_wall / args.steps, 3),
        "timed_same_result": same_timed,
        "mpx_per_s": round(fps * 512 * 512 / 1e6, 1),
        "evaluations_per_frame": {"mean": float(np.mean(evals)), "max": int(np.max(evals))},
        "plain": {"frames_per_s": round(plain_fps, 2), "frames": npl,
                  "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 3),
                                 "p99": round(float(np.percentile(lat, 99)), 3)},
                  "same_result_as_batched": plain_same},
        "batched_vs_plain": round(fps / plain_fps, 2),
        "roofline": roofline_of(tsum, "k_pass<2> (fused bilinear warp + SSD + Jtr)",
                                "k_pass_traffic_512x512_200f_nD31.txt"),
    }
    out["roofline"]["measured"] = (f"HIP events around every fused pass of {args.steps} more steps of the "
                                   f"same workload (eager launches; the batch's parts run one after the "
                                   f"other, so each pass is timed alone)")
    # the same bytes over the step time of `value`, where the parts' rounds interleave (one
    # part's pass fills the GPU while another's tail drains): a whole-step rate that includes
    # the trust-region updates and host polls
    step_gbs = tsum["pass_bytes"] / args.steps / (ms_per_step * 1e-3) / 1e9
    out["roofline"]["per_step"] = {"achieved": round(step_gbs, 1), "frac": round(step_gbs / HBM_PEAK_GBS, 4),
                                   "what": "algorithmic pass bytes of a step over ms_per_step (parts interleaved; "
                                           "updates and host polls included)"}
    if iso_ms:
        out["roofline"]["isolated"] = {"avg_launch_us": round(1e3 * iso_ms, 2), "bytes_per_launch": iso_bytes,
                                       "achieved": round(iso_bytes / (iso_ms * 1e-3) / 1e9, 1),
                                       "what": "one pass with every frame active, nothing else on the GPU"}
    batch.close()
    for f in frames:
        f.release()
    if args.end_to_end:
        out["end_to_end"] = end_to_end(F, args, ctx, st, tr)
    if args.large_frames:
        out["large_working_set"] = large_working_set(F, args, ref, st["frames"], tuning, tr, xs)

    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        host = host_cpu(args.cpu_threads)
        ncpu = min(args.cpu_frames, n)
        fw = np.stack([synth.with_offset(f) for f in st["frames"][:ncpu]])
        # parity: the checker build (no fast-math, block-tridiagonal solves) on the same frames
        x_o, s_o, _ = O.Oracle().set_reference(ref_w, nD).align_many(fw, threads=host["threads"], **tr)
        err = max(float(np.max(np.abs(xs[i] - x_o[i]))) for i in range(ncpu))
        # baseline: the Release-flag build with the reference's dense solves
        x_f, s_f, secs, wall = cpu_time(O, ref_w, nD, fw, host["threads"], tr)
        out["cpu_baseline"] = {"value": round(ncpu / wall, 3), "unit": "frames/s",
                               "cores": host["threads"], "kind": "port",
                               "sample": f"first {ncpu} of the {n} frames, one frame per thread, "
                                         f"{wall:.1f}s wall, {secs.sum():.1f}s CPU",
                               "host": host, "build": CPU_BASELINE_BUILD, "solve": CPU_BASELINE_SOLVE,
                               "mpx_per_s": round(ncpu * 512 * 512 / wall / 1e6, 2),
                               "latency_ms": {"p50": round(1e3 * float(np.percentile(secs, 50)), 1),
                                              "p99": round(1e3 * float(np.percentile(secs, 99)), 1)},
                               "evaluations_per_frame": {"mean": float(np.mean([d["evaluations"] for d in s_f])),
                                                         "max": int(np.max([d["evaluations"] for d in s_f]))},
                               "max_disp_px_vs_checker_build": round(max(float(np.max(np.abs(x_f[i] - x_o[i])))
                                                                         for i in range(ncpu)), 6)}
        out["max_disp_err_px_vs_oracle"] = err
        out["c1_beside"] = c1_beside(F, args, ctx, O, host)
    return out


def c1_beside(F, args, ctx, O, host):
    """The 256x256 float32, 20-frame, 10-resampling-point workload (nD = 11), the reference's
    CPU-warper case: the GPU's batched frames/s and the CPU baseline on the same frames."""
    from fa_amd import synth
    cfg = synth.CONFIGS["2d_256_f32"]
    st = synth.make_stack(cfg["shape"], cfg["n_frames"], cfg["nD"], seed=args.seed + 1, n_blobs=cfg["n_blobs"],
                          dtype="float32")
    ref_w = synth.with_offset(st["reference"])
    fw = np.stack([synth.with_offset(f) for f in st["frames"]])
    tr = dict(rel_delta=1e-4, max_iter=50, skip_repeat=1)
    ref = F.Reference(ctx, ref_w, cfg["nD"])
    frames = [ref.register(f) for f in fw]
    b = F.Batch(frames, **tr)
    for _ in range(max(1, args.warmup)):
        b.run()
    reps = 20
    t = time.perf_counter()
    for _ in range(reps):
        xs, _, _ = b.run()
    wall = (time.perf_counter() - t) / reps
    b.close()
    for f in frames:
        f.release()
    ref.release()
    x_o, _, _ = O.Oracle().set_reference(ref_w, cfg["nD"]).align_many(fw, threads=host["threads"], **tr)
    _, _, secs, wall_c = cpu_time(O, ref_w, cfg["nD"], fw, host["threads"], tr)
    n = len(fw)
    return {"workload": "256x256 float32, 20 frames, 10 resampling points (nD=11), TR to 1e-4 rel or 50 "
                        "iterations, batched",
            "gpu": {"frames_per_s": round(n / wall, 1), "ms_per_step": round(1e3 * wall, 3), "steps": reps},
            "cpu_baseline": {"value": round(n / wall_c, 2), "unit": "frames/s", "cores": host["threads"],
                             "kind": "port", "sample": f"all {n} frames, one frame per thread, {wall_c:.2f}s wall",
                             "latency_ms": {"p50": round(1e3 * float(np.percentile(secs, 50)), 2),
                                            "p99": round(1e3 * float(np.percentile(secs, 99)), 2)}},
            "max_disp_err_px_vs_oracle": round(max(float(np.max(np.abs(xs[i] - x_o[i]))) for i in range(n)), 6)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--frames", type=int, default=200)
    ap.add_argument("--plain-frames", type=int, default=40)
    ap.add_argument("--no-end-to-end", dest="end_to_end", action="store_false",
                    help="skip the aligner front-end run of the headline workload")
    ap.add_argument("--large-frames", type=int, default=1000,
                    help="frames of the headline workload's run beyond the last-level cache (0: skip)")
    ap.add_argument("--cpu-frames", type=int, default=200)
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="CPU baseline threads (0: hardware_concurrency within one GPU's share of the host)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--chunk", type=int, default=0)
    ap.add_argument("--seed", type=int, default=20260)
    ap.add_argument("--serve-rate", type=float, default=2000.0)
    ap.add_argument("--serve-seconds", type=float, default=2.0)
    ap.add_argument("--serve-slots", type=int, default=128)
    ap.add_argument("--serve-rounds", type=int, default=3)
    ap.add_argument("--serve-loop", choices=["native", "python"], default="native",
                    help="request loop of the serving sections: in the library (fa_server_run_stream) or "
                         "in Python over fa_server_submit_host / fa_server_step")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve512-high", type=float, default=20000.0,
                    help="second offered rate (frames/s) of the 512^2 serving run")
    ap.add_argument("--serve512-sat", type=float, default=100000.0,
                    help="third offered rate (frames/s) of the 512^2 serving run, above what it sustains (0: skip)")
    ap.add_argument("--plain-fps", type=float, default=1400.0,
                    help="plain-path frames/s used for the 512^2 serving rate when the batch section is skipped")
    ap.add_argument("--serve-lower", default="1500,1000",
                    help="lower offered rates (frames/s) the serving section also runs")
    ap.add_argument("--no-tune", action="store_true", help="default launch shape, no autotuning")
    ap.add_argument("--set-tuning", default="",
                    help="launch shape key=value,... applied instead of autotuning (e.g. the shape a tuned "
                         "run reported), for profiling the same launches")
    ap.add_argument("--no-isolated", action="store_true", help="skip the isolated one-pass timing")
    ap.add_argument("--sections", default="batch,serve,precision,uneven",
                    help="comma list of: batch (the headline 512^2 workload), serve, precision (3-D), "
                         "uneven (64 stacks)")
    ap.add_argument("--c3-frames", type=int, default=50)
    ap.add_argument("--c3-oracle", type=int, default=8, help="volumes checked against the oracles")
    ap.add_argument("--c4-stacks", type=int, default=64)
    ap.add_argument("--tune-frames", type=int, default=0, help="frames the chunk-size tuning uses (0: all)")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="MiB copied device-to-device once (calibrates rocprof byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (frames are "
              f"independent given the reference: several GPUs would each take whole frames)", file=sys.stderr)
        sys.exit(2)

    secs = set(args.sections.split(","))
    import fa_amd as F
    ctx = F.Context(0)
    if "batch" in secs:
        out = headline(F, args, ctx)
    else:
        out = {"metric": METRIC, "value": None, "unit": "frames/s", "n_gpus": 1, "steps": args.steps,
               "warmup": args.warmup, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
               "dtype": "f32", "data": "synthetic", "config": {"workload": "sections " + args.sections}}

    if args.copy_probe:
        out["copy_probe_mib"] = copy_probe(args.copy_probe)

    if "serve" in secs and not args.no_serve:
        plain = out.get("plain", {}).get("frames_per_s") or args.plain_fps
        out["serving_512"] = serving_512(F, args, ctx, plain)
        out["serving"] = serving(F, args, ctx)
        log(f"serving: {out['serving']['frames_per_s']} frames/s, p50 {out['serving']['latency_ms']['p50']} ms, "
            f"p99 {out['serving']['latency_ms']['p99']} ms")
    if "precision" in secs:
        out["precision_3d"] = precision_3d(F, args, ctx)
        log(f"precision_3d: {out['precision_3d']['volumes_per_s']} volumes/s")
    if "uneven" in secs:
        out["batched_uneven"] = uneven_batch(F, args, ctx)
        log(f"batched_uneven: {out['batched_uneven']['frames_per_s']} frames/s")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
