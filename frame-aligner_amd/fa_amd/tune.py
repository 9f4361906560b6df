"""Autotuner of the batched trust-region runs.

Two levels, both timed on the caller's real workload (no model of the hardware):

* per reference, the work-item size `chunk_px`: the tile of pixels one workgroup of the fused
  pass takes (one node interval inside one plane, at most chunk_px pixels); it fixes the LDS
  tile the workgroup stages and the number of workgroups per pass;
* per batch, the launch shape of fa_tuning: the LDS staging budget (or none), the grid cap in
  workgroups per compute unit, trust-region rounds per host poll, hipGraph replay of those
  rounds, and whether the trust-region update keeps its factor blocks in LDS.

The objective is the wall time of whole Batch.run() calls (untimed, so graph replay is
allowed), median of `reps`. The search is coordinate descent: one knob at a time over its
candidate values from the current best, repeated `sweeps` times. Every candidate gives the same
results (tuning changes where and when work runs, never what it computes), so the tuner can
run on the production frames themselves. Results can be kept in a JSON cache keyed by the
workload's signature (sizes, nodes, dtypes, frame count, device).
"""
import json
import os
import time

import numpy as np

# Candidate values of each fa_tuning knob, in the order coordinate descent visits them.
SPACE = {
    "blocks_per_cu": [8, 16, 32, 64],
    "check_every": [4, 8, 16, 32],
    "use_graph": [1, 0],
    "use_lds": [1, 0],
    "tr_lds": [1, 0],
}
CHUNKS = [4096, 2048, 8192, 16384]  # the library's default first


def signature(frames, ctx=None):
    """Workload key: per frame kind (shape, nodes, dtype) the count, plus the device."""
    kinds = {}
    for f in frames:
        k = (tuple(f.ref.shape), f.ref.nD, f.dtype)
        kinds[k] = kinds.get(k, 0) + 1
    dev = ctx.info() if ctx is not None else {}
    parts = [f"{'x'.join(map(str, s))}/{nD}/{'u16' if dt == 1 else 'f32'}*{n}"
             for (s, nD, dt), n in sorted(kinds.items())]
    return ";".join(parts) + f"@cu{dev.get('n_cu', '?')}"


def time_run(batch, reps=3, x0=None):
    """Median wall seconds of batch.run() (one warm-up run first)."""
    batch.run(x0)
    ts = []
    for _ in range(reps):
        t = time.perf_counter()
        batch.run(x0)
        ts.append(time.perf_counter() - t)
    return float(np.median(ts))


def tune_batch(batch, reps=3, sweeps=2, space=None, log=None, x0=None, gain=0.02):
    """Coordinate descent over fa_tuning: per knob, every candidate value is timed from the
    current best, and the fastest (the first in the knob's order on a tie) replaces the current
    best when it is faster by more than `gain` (relative), so timing noise does not move the
    shape. Leaves the batch at the best shape and returns (best tuning dict, best seconds per
    run, list of (tuning, seconds) trials)."""
    space = space or SPACE
    was_timing = getattr(batch, "_timing", False)
    batch.set_timing(False)
    best = dict(batch.tuning)
    batch.set_tuning(**best)
    best_t = time_run(batch, reps, x0)
    trials = [(dict(best), best_t)]
    for _ in range(sweeps):
        improved = False
        for knob, values in space.items():
            top, top_t = None, None
            for v in values:
                if best.get(knob) == v:
                    continue
                cand = dict(best, **{knob: v})
                batch.set_tuning(**cand)
                t = time_run(batch, reps, x0)
                trials.append((dict(cand), t))
                if log:
                    log(f"tune {knob}={v}: {t * 1e3:.3f} ms (best {best_t * 1e3:.3f})")
                if top_t is None or t < top_t:
                    top, top_t = cand, t
            if top is not None and top_t < best_t * (1.0 - gain):
                best, best_t, improved = top, top_t, True
            batch.set_tuning(**best)
        if not improved:
            break
    batch.set_tuning(**best)
    batch.set_timing(was_timing)
    return best, best_t, trials


def tune_chunk(ctx, reference, nD, raw_frames, chunks=None, reps=3, tr=None, log=None, gain=0.02,
               **ref_kw):
    """Work-item size for one acquisition: for each candidate chunk_px, builds the reference,
    registers the frames, tunes the batch briefly and times it. The first candidate (the
    library's default size) is kept unless another is faster by more than `gain`. Returns
    (best chunk_px, {chunk_px: seconds})."""
    from . import Batch, Reference
    tr = tr or dict(rel_delta=1e-4, max_iter=50, skip_repeat=1)
    res = {}
    for c in chunks or CHUNKS:
        ref = Reference(ctx, reference, nD, chunk_px=c, **ref_kw)
        frames = [ref.register(f) for f in raw_frames]
        b = Batch(frames, **tr)
        _, t, _ = tune_batch(b, reps=reps, sweeps=1, space={"blocks_per_cu": SPACE["blocks_per_cu"]})
        res[c] = t
        if log:
            log(f"tune chunk_px={c}: {t * 1e3:.3f} ms")
        b.close()
        for f in frames:
            f.release()
        ref.release()
    first = next(iter(res))
    best = min(res, key=res.get)
    return (best if res[best] < res[first] * (1.0 - gain) else first), res


def load_cache(path):
    try:
        with open(path) as fh:
            return json.load(fh)
    except (OSError, ValueError):
        return {}


def save_cache(path, cache):
    d = os.path.dirname(path)
    if d:
        os.makedirs(d, exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as fh:
        json.dump(cache, fh, indent=1, sort_keys=True)
    os.replace(tmp, path)


def tuned(batch, ctx, cache_path=None, **kw):
    """Applies the cached shape for this workload if there is one, else tunes and (with a
    cache_path) stores it. Returns the tuning dict in use."""
    key = signature(batch.frames, ctx)
    cache = load_cache(cache_path) if cache_path else {}
    if key in cache:
        batch.set_tuning(**cache[key]["tuning"])
        return cache[key]["tuning"]
    best, t, _ = tune_batch(batch, **kw)
    if cache_path:
        cache[key] = {"tuning": best, "seconds_per_run": t}
        save_cache(cache_path, cache)
    return best
