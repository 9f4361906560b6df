"""Command front end of the aligner: the reference's MATLAB entry point
(FrameAlignerMex/FrameAlignerMex.cpp:18-162) as a Python function over fa_amd.Aligner.

    idx = frame_aligner(params, scan, n_frames)          # create; returns the aligner index
    frame_aligner(idx, "SetNumFrames", n)
    frame_aligner(idx, "SetReference", frame_t, reference)
    frame_aligner(idx, "AddFrame", frame_t, frame)        # queued; aligned asynchronously
    n = frame_aligner(idx, "NumTasksRemaining")
    img = frame_aligner(idx, "GetRealignedFrame", frame_t) # uint16; "Result not ready" until done
    frame_aligner(idx, "WriteRealignmentInfo", filename)   # FrameWarpAligner::writeRealignmentInfo
    frame_aligner(idx, "Delete")

params: dict with the reference's fields and defaults (type 2 = Warp, spatial_binning 1,
frame_binning 1, n_resampling_points 10, smoothing 0, correlation_threshold 0,
coverage_threshold 0); scan: line_duration (1), interline_duration (1), n_x, n_y, n_z (1),
bidirectional (0), and interframe_duration, which the reference fixes at 1 (its planes then
overlap in time and its node map is not monotonic; this library rejects such a map), here
defaulting to one frame duration.

Where the reference pushes each AddFrame onto a thread pool task (one frame per task,
FrameAlignerMex.cpp:110-130), each aligner here has one worker thread that takes every frame
queued so far and aligns them as one batch on the GPU (fa_aligner_add_frames): frames that
arrive while a batch runs wait for the next one. Results (D, the realigned uint16 frame,
correlations, coverage) are kept in host memory; the reference's disk-backed Cache is out of
scope. AddFrame before SetReference raises the reference's "Reference imaging not yet set" at
once instead of inside the worker.
"""
import threading

import numpy as np

from . import Aligner, Context, FaError, write_realignment_info

_lock = threading.Lock()
_aligners = []
_ctx = None

PARAM_DEFAULTS = dict(type=2, spatial_binning=1, frame_binning=1, n_resampling_points=10, smoothing=0.0,
                      correlation_threshold=0.0, coverage_threshold=0.0)
SCAN_DEFAULTS = dict(line_duration=1.0, interline_duration=1.0, n_z=1, bidirectional=0)


class _Slot:
    """One aligner of the registry: parameters, the device aligner (after SetReference), the
    queue of frames, the worker and the results."""

    def __init__(self, params, scan, n_frames):
        self.params = params
        self.scan = scan
        self.n_frames = n_frames
        self.aligner = None
        self.queue = []
        self.pending = 0          # queued + in the running batch
        self.results = {}
        self.cv = threading.Condition()
        self.worker = None
        self.closed = False
        self.error = None

    def _run(self):
        while True:
            with self.cv:
                while not self.queue and not self.closed:
                    self.cv.wait()
                if self.closed and not self.queue:
                    return
                batch, self.queue = self.queue, []
            ts = [t for t, _ in batch]
            frames = [f for _, f in batch]
            try:
                D, res, imgs = self.aligner.add_frames(frames, outputs=("realigned",))
                done = {t: dict(D=D[i], realigned=imgs["realigned"][i], **res[i]) for i, t in enumerate(ts)}
            except Exception as e:  # the reference swallows a task's exception (FrameAlignerMex.cpp:118-127)
                done = {}
                self.error = e
            with self.cv:
                self.results.update(done)
                self.pending -= len(batch)
                self.cv.notify_all()

    def add(self, frame_t, frame):
        if self.aligner is None:
            raise FaError(-6, "Reference imaging not yet set")
        with self.cv:
            self.queue.append((frame_t, np.array(frame, copy=True)))
            self.pending += 1
            if self.worker is None:
                self.worker = threading.Thread(target=self._run, daemon=True)
                self.worker.start()
            self.cv.notify_all()

    def wait(self):
        with self.cv:
            while self.pending:
                self.cv.wait()

    def close(self):
        with self.cv:
            self.closed = True
            self.cv.notify_all()
        if self.worker is not None:
            self.worker.join()
        if self.aligner is not None:
            self.aligner.close()


def _context():
    global _ctx
    if _ctx is None:
        _ctx = Context(0)
    return _ctx


def _create(params, scan, n_frames):
    p = dict(PARAM_DEFAULTS, **(params or {}))
    sc = dict(SCAN_DEFAULTS, **(scan or {}))
    if p["frame_binning"] < 1 or p["spatial_binning"] < 1:
        raise FaError(-1, "Input condition not met")
    for k in ("n_x", "n_y"):
        if k not in sc:
            raise FaError(-1, f"missing scan field {k}")
    slot = _Slot(p, sc, int(n_frames))
    with _lock:
        # the first empty place (FrameAlignerMex.cpp:63-74)
        for i, a in enumerate(_aligners):
            if a is None:
                _aligners[i] = slot
                return i
        _aligners.append(slot)
        return len(_aligners) - 1


def _slot(idx):
    with _lock:
        if not isinstance(idx, (int, np.integer)) or idx < 0 or idx >= len(_aligners) or _aligners[idx] is None:
            raise FaError(-1, "Invalid aligner index specified")
        return _aligners[idx]


def frame_aligner(*args):
    """Dispatches one command of the reference's MEX interface (see the module docstring)."""
    if len(args) == 3 and isinstance(args[0], dict) and isinstance(args[1], dict):
        return _create(*args)
    if len(args) < 2:
        raise FaError(-1, "Input condition not met")
    idx, command = args[0], args[1]
    if not isinstance(command, str):
        raise FaError(-1, "Second argument should be a command string")
    s = _slot(idx)
    if command == "SetNumFrames":
        if len(args) < 3 or int(args[2]) <= 0:
            raise FaError(-1, "Input condition not met")
        s.n_frames = int(args[2])
    elif command == "SetReference":
        if len(args) < 4:
            raise FaError(-1, "Input condition not met")
        ref = np.asarray(args[3])
        sc, p = s.scan, s.params
        ny = int(sc["n_y"])
        interline = float(sc["interline_duration"])
        line = float(sc["line_duration"])
        interframe = float(sc.get("interframe_duration", ny * max(interline, line)))
        if s.aligner is not None:
            s.wait()
            s.aligner.close()
        shape = (int(sc["n_z"]), ny, int(sc["n_x"]))
        s.aligner = Aligner(_context(), ref.reshape(shape), type=int(p["type"]),
                            n_resampling_points=int(p["n_resampling_points"]), smoothing=float(p["smoothing"]),
                            spatial_binning=int(p["spatial_binning"]), line=line, interline=interline,
                            interframe=interframe)
    elif command == "AddFrame":
        if len(args) < 4:
            raise FaError(-1, "Input condition not met")
        s.add(int(args[2]), np.asarray(args[3]).reshape(s.aligner.shape if s.aligner else -1))
    elif command == "NumTasksRemaining":
        with s.cv:
            return s.pending
    elif command == "GetRealignedFrame":
        if len(args) < 3:
            raise FaError(-1, "Input condition not met")
        with s.cv:
            r = s.results.get(int(args[2]))
        if r is None:
            raise FaError(-6, "Result not ready")
        img = r["realigned"]
        return img[0] if img.shape[0] == 1 else img   # 2-D frames come back 2-D (output2d)
    elif command == "GetRealignmentResult":
        with s.cv:
            r = s.results.get(int(args[2]))
        if r is None:
            raise FaError(-6, "Result not ready")
        return dict(r)
    elif command == "WriteRealignmentInfo":
        s.wait()
        with s.cv:
            ts = sorted(s.results)
            rs = [s.results[t] for t in ts]
        if rs:
            sc = s.scan
            write_realignment_info(str(args[2]), np.stack([r["D"] for r in rs]),
                                   [r["unaligned_correlation"] for r in rs], [r["correlation"] for r in rs],
                                   [r["coverage"] for r in rs], n_x=int(sc["n_x"]), n_y=int(sc["n_y"]),
                                   n_z=int(sc["n_z"]), line=float(sc["line_duration"]),
                                   interline=float(sc["interline_duration"]),
                                   interframe=float(sc.get("interframe_duration",
                                                           int(sc["n_y"]) * float(sc["interline_duration"]))))
    elif command == "Delete":
        with _lock:
            _aligners[idx] = None
        s.close()
    else:
        raise FaError(-1, f"Unknown command {command}")
    return None
