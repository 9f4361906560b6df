"""Serving loop: frames streaming from several acquisitions, aligned by continuous batching.

Every acquisition keeps its reference (packed intensity + gradients, node map, Hessian
factors) resident in HBM. Frames arrive with exponential inter-arrival times and are handed
to a Server, which admits a frame when one of its slots and room for the frame in its device
frame arena are free (admission by free memory), uploads it on a copy stream while the
frames already admitted compute, and runs the trust-region rounds of all admitted frames
together: a frame that arrives while others are on the GPU joins their launches and leaves
when it finishes. Latency is arrival -> result on the host.

The reference aligner has no serving loop: its MEX front end queues addFrame calls on a
thread pool (FrameAlignerMex/FrameAlignerMex.cpp:110-130), one frame per task.
"""
import time

import numpy as np

from . import synth

# Acquisition sizes (z, y, x) of the serving workload, 256^2 x 1 up to 1024^2 x 8.
SERVE_SHAPES = [(1, 256, 256), (1, 512, 512), (1, 1024, 1024), (4, 256, 256),
                (4, 512, 512), (8, 256, 256), (8, 512, 512), (8, 1024, 1024)]


def make_acquisitions(seed=2026, shapes=SERVE_SHAPES, pool=3, node_range=(10, 40), mean=50.0):
    """One seeded acquisition per shape: uint16 Poisson blob phantom (blob count scaled with
    the imaged area and depth), resampling points per plane drawn uniformly from node_range
    (nD = points * planes + 1, as the aligner sets it, FrameWarpAligner.cpp:100), `pool`
    distinct frames that the arrival stream cycles through. Returns a list of dicts
    (synth.make_stack's keys)."""
    rng = np.random.default_rng(seed)
    out = []
    for i, shape in enumerate(shapes):
        nz, ny, nx = shape
        nD = int(rng.integers(node_range[0], node_range[1] + 1)) * nz + 1
        blobs = int(600 * ny * nx / 512 ** 2 * max(1.0, nz / 4))
        st = synth.make_stack(shape, pool, nD, seed=seed * 100 + i, n_blobs=blobs,
                              sigma=(2.0, 6.0) if nz == 1 else (1.5, 4.0), mean=mean,
                              dtype="uint16")
        out.append(st)
    return out


def arrivals(n_acq, rate, duration, seed):
    """Merged Poisson stream at `rate` frames/s over `duration` s, each arrival from an
    acquisition drawn uniformly (so each acquisition streams at rate / n_acq).
    Returns (times, acquisition index) arrays."""
    rng = np.random.default_rng(seed)
    n = max(1, int(rate * duration * 1.5) + 16)
    t = np.cumsum(rng.exponential(1.0 / rate, n))
    t = t[t < duration]
    return t, rng.integers(0, n_acq, t.size)


class ServeLoop:
    """The request loop. refs: one fa_amd.Reference per acquisition; pools: per acquisition,
    the raw uint16 frames its stream cycles through. The pools are copied once into pinned
    host memory (where a frame grabber would DMA them), so uploads are asynchronous."""

    def __init__(self, server, refs, pools, rounds=3):
        from . import PinnedArray
        self.server = server
        self.refs = refs
        self.pinned = [PinnedArray(p.shape, p.dtype) for p in pools]
        for dst, src in zip(self.pinned, pools):
            dst.array[...] = src
        self.rounds = rounds

    def close(self):
        for p in self.pinned:
            p.close()

    def run(self, times, acq, keep=None, native=False):
        """Streams the arrivals in real time. Returns dict with per-frame latency (s), the
        results of the frames whose index is in `keep`, and loop counters. native: the loop runs
        in the library (fa_server_run_stream) instead of here."""
        if native:
            return self._run_native(times, acq, keep)
        S = self.server
        n = len(times)
        lat = np.full(n, np.nan)
        evals = np.zeros(n, np.int64)
        kept = {}
        nxt = 0
        done = 0
        steps = 0
        max_flight = 0
        max_queued = 0
        busy = 0
        t0 = time.perf_counter()
        while done < n:
            now = time.perf_counter() - t0
            while nxt < n and times[nxt] <= now:
                a = int(acq[nxt])
                pool = self.pinned[a].array
                S.submit_host(self.refs[a], pool[nxt % len(pool)], tag=nxt)
                nxt += 1
                busy += 1
            if busy == 0:
                time.sleep(max(0.0, times[nxt] - (time.perf_counter() - t0)))
                continue
            in_flight, queued, res = S.step(self.rounds)
            steps += 1
            max_flight = max(max_flight, in_flight)
            max_queued = max(max_queued, queued)
            tnow = time.perf_counter() - t0
            for tag, x, st in res:
                lat[tag] = tnow - times[tag]
                evals[tag] = st["evaluations"]
                if keep is not None and tag in keep:
                    kept[tag] = (x, st)
                done += 1
                busy -= 1
        wall = time.perf_counter() - t0
        return dict(latency=lat, kept=kept, wall=wall, steps=steps, max_in_flight=max_flight,
                    max_queued=max_queued, evaluations=evals)

    def _run_native(self, times, acq, keep=None):
        n = len(times)
        refs, frames = [], []
        for i in range(n):
            a = int(acq[i])
            pool = self.pinned[a].array
            refs.append(self.refs[a])
            frames.append(pool[i % len(pool)])
        lat, res, cnt = self.server.run_stream(times, refs, frames, rounds=self.rounds)
        kept = {i: res[i] for i in range(n) if keep is not None and i in keep}
        evals = np.array([r[1]["evaluations"] for r in res], np.int64)
        return dict(latency=lat, kept=kept, wall=cnt["wall_s"], steps=cnt["steps"],
                    max_in_flight=cnt["max_in_flight"], max_queued=cnt["max_queued"], evaluations=evals)


def summarize(res, times, acq, pools):
    lat = res["latency"] * 1e3
    n = len(times)
    px = sum(pools[int(a)][0].size for a in acq)
    ev = res.get("evaluations")
    extra = {}
    if ev is not None:
        extra["evaluations_per_frame"] = {"mean": round(float(np.mean(ev)), 2), "max": int(np.max(ev))}
        per = {}
        for a, e, l in zip(acq, ev, lat):
            per.setdefault(int(a), []).append((e, l))
        extra["per_acquisition"] = {int(a): {"frames": len(v), "evaluations_mean": round(float(np.mean([e for e, _ in v])), 1),
                                             "latency_p50_ms": round(float(np.percentile([l for _, l in v], 50)), 2),
                                             "latency_p99_ms": round(float(np.percentile([l for _, l in v], 99)), 2)}
                                    for a, v in sorted(per.items())}
    return {"frames": n, "wall_s": round(res["wall"], 3), **extra,
            "frames_per_s": round(n / res["wall"], 1),
            "offered_frames_per_s": round(n / max(times[-1], 1e-9), 1),
            "mpx_per_s": round(px / res["wall"] / 1e6, 1),
            "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 3),
                           "p99": round(float(np.percentile(lat, 99)), 3),
                           "max": round(float(lat.max()), 3)},
            "steps": res["steps"], "max_in_flight": res["max_in_flight"],
            "max_queued": res["max_queued"]}
