"""Python front end of the gfx950 warp-registration library (libfa_hip.so).

Mirrors the reference's warper interface (FrameAligner/FrameWarper.h:95-143,
GpuFrameWarper.h:31-72): setReference, registerFrame / deregisterFrame,
getError, getJacobian, warpImageInterpolated, warpPoint, H; plus the device
trust-region loop of FrameWarpAligner::addFrame (FrameWarpAligner.cpp:200-215)
for one frame (align) or many (Batch).

Everything runs through the C ABI of include/frame_aligner.h. There is no CPU
fallback: if the HIP library is missing or no gfx950 GPU is present, the
calls raise.
"""
import ctypes as C
import os
import weakref

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
LIB_PATH = os.path.join(PKG, "lib", "libfa_hip.so")

FA_F32, FA_U16 = 0, 1
STOP_NAMES = {1: "abs_delta", 2: "rel_delta", 3: "max_iter", 4: "no_progress", 5: "nonfinite",
              6: "radius", 7: "max_evals"}


# Fields of a trust-region trace record (fa_batch_set_trace), then x, g and p.
TRACE_FIELDS = ["evals", "f_new", "f", "pred", "meas", "rho", "accept", "radius", "iterations",
                "status", "lambda", "sub_iters", "pnorm", "gnorm", "gp", "rejected"]


class FaError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{msg} (code {code})")
        self.code = code


class ScanParams(C.Structure):
    """ImageScanParameters constructor arguments (ImageScanParameters.h:7-20)."""
    _fields_ = [("line_duration", C.c_double), ("interline_duration", C.c_double),
                ("interframe_duration", C.c_double), ("n_x", C.c_int32), ("n_y", C.c_int32),
                ("n_z", C.c_int32), ("bidirectional", C.c_int32)]


class TrParams(C.Structure):
    _fields_ = [("radius", C.c_double), ("min_delta", C.c_double), ("rel_delta", C.c_double),
                ("max_iter", C.c_int32), ("max_evals", C.c_int32), ("sub_eps", C.c_double),
                ("sub_max_iter", C.c_int32), ("skip_repeat", C.c_int32)]


class AlignStats(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("evaluations", C.c_int32), ("rejected", C.c_int32),
                ("status", C.c_int32), ("f_final", C.c_double), ("radius_final", C.c_double)]


class Tuning(C.Structure):
    """Launch shape of the fused pass (fa_tuning in include/frame_aligner.h)."""
    _fields_ = [("lds_floats", C.c_int32), ("blocks_per_cu", C.c_int32), ("use_lds", C.c_int32),
                ("check_every", C.c_int32), ("use_graph", C.c_int32), ("tr_lds", C.c_int32)]


class RealignParams(C.Structure):
    """RealignmentParameters (AbstractFrameAligner.h:31-48), the fields addFrame reads."""
    _fields_ = [("type", C.c_int32), ("spatial_binning", C.c_int32), ("n_resampling_points", C.c_int32),
                ("pad", C.c_int32), ("smoothing", C.c_double)]


class AddResult(C.Structure):
    _fields_ = [("start", C.c_int32), ("pad", C.c_int32), ("rigid_shift", C.c_double * 3),
                ("start_error", C.c_double * 3), ("stats", AlignStats), ("correlation", C.c_double),
                ("unaligned_correlation", C.c_double), ("coverage", C.c_double)]


class AddOutputs(C.Structure):
    _fields_ = [("realigned", C.c_void_p), ("preserving", C.c_void_p), ("mask", C.c_void_p),
                ("warped", C.c_void_p)]


REALIGN_TYPES = {"none": 0, "translation": 1, "warp": 2}


class ServerResult(C.Structure):
    _fields_ = [("tag", C.c_int64), ("stats", AlignStats)]


class StreamCounts(C.Structure):
    _fields_ = [("steps", C.c_int64), ("max_in_flight", C.c_int64), ("max_queued", C.c_int64),
                ("wall_s", C.c_double)]


class BatchTiming(C.Structure):
    _fields_ = [("wall_ms", C.c_double), ("pass_ms", C.c_double), ("pass_launches", C.c_int64),
                ("pass_bytes", C.c_double), ("tr_rounds", C.c_int64), ("pass_px", C.c_double)]


EXPORTS = [
    "fa_last_error", "fa_tr_default", "fa_create", "fa_destroy", "fa_set_reference",
    "fa_reference_release", "fa_reference_info", "fa_get_hessian", "fa_get_node_map",
    "fa_get_gradient", "fa_register_frame", "fa_deregister_frame", "fa_clear_temp", "fa_get_error",
    "fa_get_jacobian", "fa_warp_interpolated", "fa_warp_image", "fa_warp_point", "fa_align", "fa_align_noalloc",
    "fa_batch_create",
    "fa_batch_destroy", "fa_batch_set_timing", "fa_batch_set_trace", "fa_batch_get_trace",
    "fa_batch_get_tuning", "fa_batch_set_tuning",
    "fa_batch_run", "fa_batch_eval", "fa_align_batch",
    "fa_context_stream", "fa_context_mem_info", "fa_context_info", "fa_server_create", "fa_server_destroy",
    "fa_server_submit", "fa_server_submit_host", "fa_server_step", "fa_server_results", "fa_server_run_stream",
    "fa_server_memory", "fa_host_alloc", "fa_host_free",
    "fa_aligner_create", "fa_aligner_destroy", "fa_aligner_reference", "fa_aligner_preprocess",
    "fa_aligner_rigid_shift", "fa_aligner_add_frames", "fa_aligner_clear_temp", "fa_aligner_realign_as_frame",
    "fa_shift_pixel",
    "fa_write_realignment_info",
]

_lib = None


def load_library(path=None):
    """Loads libfa_hip.so (built by frame-aligner_amd/build.sh); raises if absent."""
    global _lib
    if _lib is not None:
        return _lib
    path = path or LIB_PATH
    if not os.path.exists(path):
        raise FaError(-4, f"HIP library not built: {path} (run frame-aligner_amd/build.sh)")
    L = C.CDLL(path)
    P = C.c_void_p
    i32 = C.c_int32
    L.fa_last_error.restype = C.c_char_p
    L.fa_tr_default.argtypes = [C.POINTER(TrParams)]
    L.fa_tr_default.restype = None
    L.fa_create.argtypes = [C.c_int, C.POINTER(P)]
    L.fa_destroy.argtypes = [P]
    L.fa_set_reference.argtypes = [P, P, i32, i32, i32, i32, C.POINTER(ScanParams), i32, C.POINTER(P)]
    L.fa_reference_release.argtypes = [P]
    L.fa_reference_info.argtypes = [P, P]
    L.fa_get_hessian.argtypes = [P, P]
    L.fa_get_node_map.argtypes = [P, P, P, P]
    L.fa_get_gradient.argtypes = [P, P, P, P]
    L.fa_register_frame.argtypes = [P, P, i32, C.c_float, C.POINTER(P)]
    L.fa_deregister_frame.argtypes = [P]
    L.fa_clear_temp.argtypes = [P]
    L.fa_get_error.argtypes = [P, P, P]
    L.fa_get_jacobian.argtypes = [P, P, P, P]
    L.fa_warp_interpolated.argtypes = [P, P, P, C.c_float]
    L.fa_warp_point.argtypes = [P, P, i32, i32, i32, i32, P]
    L.fa_warp_image.argtypes = [P, P, P, P, i32]
    L.fa_align.argtypes = [P, C.POINTER(TrParams), P, C.POINTER(AlignStats)]
    L.fa_align_noalloc.argtypes = [P, C.POINTER(TrParams), P, C.POINTER(AlignStats)]
    L.fa_batch_create.argtypes = [P, i32, C.POINTER(TrParams), C.POINTER(P)]
    L.fa_batch_destroy.argtypes = [P]
    L.fa_batch_set_timing.argtypes = [P, i32]
    L.fa_batch_set_trace.argtypes = [P, i32]
    L.fa_batch_get_trace.argtypes = [P, P]
    L.fa_batch_get_tuning.argtypes = [P, C.POINTER(Tuning)]
    L.fa_batch_set_tuning.argtypes = [P, C.POINTER(Tuning)]
    L.fa_batch_run.argtypes = [P, P, P, C.POINTER(BatchTiming)]
    L.fa_batch_eval.argtypes = [P, P, P, P]
    L.fa_align_batch.argtypes = [P, i32, C.POINTER(TrParams), P, P]
    L.fa_context_mem_info.argtypes = [P, P, P]
    L.fa_context_info.argtypes = [P, P]
    L.fa_aligner_create.argtypes = [P, C.POINTER(RealignParams), C.POINTER(ScanParams), P, i32, C.POINTER(P)]
    L.fa_aligner_destroy.argtypes = [P]
    L.fa_aligner_reference.argtypes = [P, C.POINTER(P), C.POINTER(i32)]
    L.fa_aligner_preprocess.argtypes = [P, P, i32, P]
    L.fa_aligner_rigid_shift.argtypes = [P, P, i32, P]
    L.fa_aligner_add_frames.argtypes = [P, P, i32, i32, C.POINTER(TrParams), P, C.POINTER(AddOutputs), P]
    L.fa_aligner_clear_temp.argtypes = [P]
    L.fa_aligner_realign_as_frame.argtypes = [P, P, P, i32, i32, i32, P]
    L.fa_shift_pixel.argtypes = [P, P, P]
    L.fa_write_realignment_info.argtypes = [C.c_char_p, C.POINTER(ScanParams), i32, i32, P, P, P, P]
    L.fa_server_create.argtypes = [P, P, i32, i32, C.c_uint64, C.POINTER(TrParams), C.POINTER(P)]
    L.fa_server_submit_host.argtypes = [P, P, P, i32, C.c_float, P, C.c_int64]
    L.fa_server_memory.argtypes = [P, P, P]
    L.fa_host_alloc.argtypes = [C.c_uint64, C.POINTER(P)]
    L.fa_host_free.argtypes = [P]
    L.fa_server_destroy.argtypes = [P]
    L.fa_server_submit.argtypes = [P, P, P, C.c_int64]
    L.fa_server_step.argtypes = [P, i32, P, P]
    L.fa_server_results.argtypes = [P, P, i32, P]
    L.fa_server_run_stream.argtypes = [P, C.c_int64, P, P, P, P, P, P, P, i32, P, P, P]
    L.fa_context_stream.argtypes = [P]
    L.fa_context_stream.restype = P
    _lib = L
    return L


def _check(rc):
    if rc != 0:
        raise FaError(rc, _lib.fa_last_error().decode())


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def tr_params(**kw):
    """Trust-region controls; defaults are the reference's call (radius 40, |df| < 1e-8)."""
    L = load_library()
    p = TrParams()
    L.fa_tr_default(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def stats_dict(s):
    d = {f: getattr(s, f) for f, _ in AlignStats._fields_}
    d["stop"] = STOP_NAMES.get(d["status"], "?")
    return d


class Context:
    """One GPU (GpuFrameWarper's device, GpuFrameWarper.cpp:10-44)."""

    def __init__(self, device=0):
        self.L = load_library()
        h = C.c_void_p()
        _check(self.L.fa_create(device, C.byref(h)))
        self.h = h
        self._children = weakref.WeakSet()

    def _adopt(self, obj):
        self._children.add(obj)

    def close(self):
        """Releases every batch, frame and reference made on this context, then the context."""
        if getattr(self, "h", None):
            kids = list(self._children)
            for kind in (Aligner, Server, Batch, Frame, Reference):
                for k in kids:
                    if isinstance(k, kind):
                        k._release_handle()
            self.L.fa_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def stream(self):
        return self.L.fa_context_stream(self.h)

    def mem_info(self):
        """(free, total) device bytes."""
        f, t = C.c_uint64(), C.c_uint64()
        _check(self.L.fa_context_mem_info(self.h, C.byref(f), C.byref(t)))
        return f.value, t.value

    def info(self):
        """dict(device, n_cu, lds_per_block, wave_size, lds_config)."""
        o = np.zeros(5, np.int32)
        _check(self.L.fa_context_info(self.h, _ptr(o)))
        return dict(device=int(o[0]), n_cu=int(o[1]), lds_per_block=int(o[2]), wave_size=int(o[3]),
                    lds_config=int(o[4]))


class Reference:
    """Per-acquisition state resident in HBM: reference, Sobel gradient, node tables,
    block-tridiagonal JtJ and its factor (AbstractFrameWarper::setReference)."""

    def __init__(self, ctx, reference, nD, line=1.0, interline=1.0, interframe=None, chunk_px=0):
        self.ctx = ctx
        self.L = ctx.L
        ref = np.ascontiguousarray(reference, dtype=np.float32)
        if ref.ndim == 2:
            ref = ref[None]
        nz, ny, nx = ref.shape
        if interframe is None:
            interframe = ny * max(interline, line)
        sp = ScanParams(line, interline, interframe, nx, ny, nz, 0)
        h = C.c_void_p()
        _check(self.L.fa_set_reference(ctx.h, _ptr(ref), nz, ny, nx, nD, C.byref(sp), chunk_px, C.byref(h)))
        self.h = h
        ctx._adopt(self)
        info = np.zeros(6, np.int32)
        _check(self.L.fa_reference_info(self.h, _ptr(info)))
        self.shape = (nz, ny, nx)
        self.nD = nD
        self.n_dim = int(info[4])
        self.n_chunks = int(info[5])
        self.n_px = nz * ny * nx

    def _release_handle(self):
        """Frees the device copy; raises FaError (and keeps it) while frames or servers of
        this reference are alive."""
        if getattr(self, "h", None):
            _check(self.L.fa_reference_release(self.h))
            self.h = None

    release = _release_handle

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass

    def clear_temp(self):
        """GpuFrameWarper::clearTemp: frees the working spaces cached for this reference's
        frames (rebuilt on next use); the frames stay registered."""
        _check(self.L.fa_clear_temp(self.h))

    @property
    def H(self):
        n = self.nD * self.n_dim
        H = np.zeros((n, n))
        _check(self.L.fa_get_hessian(self.h, _ptr(H)))
        return H

    def node_map(self):
        Di = np.zeros(self.n_px, np.uint16)
        Df = np.zeros(self.n_px, np.float32)
        Dr = np.zeros((self.nD, 2), np.int32)
        _check(self.L.fa_get_node_map(self.h, _ptr(Di), _ptr(Df), _ptr(Dr)))
        return Di, Df, Dr

    def gradient(self):
        g = [np.zeros(self.n_px, np.float32) for _ in range(3)]
        _check(self.L.fa_get_gradient(self.h, *[_ptr(a) for a in g]))
        return [a.reshape(self.shape) for a in g]

    def warp_point(self, D, x, y, z, spatial_binning=1):
        D = as_D(D, self.nD, self.n_dim)
        o = np.zeros(3)
        _check(self.L.fa_warp_point(self.h, _ptr(D), x, y, z, spatial_binning, _ptr(o)))
        return o

    def register(self, frame, offset=None):
        return Frame(self, frame, offset)


def as_D(D, nD, n_dim):
    """nD x 3 doubles from an nD x 3 array or a D2col vector (OptimisationModel.cpp:14-42)."""
    D = np.asarray(D, np.float64)
    if D.ndim == 1:
        x = D.reshape(nD, n_dim)
        D = np.zeros((nD, 3))
        D[:, :n_dim] = x
    return np.ascontiguousarray(D)


class Frame:
    """A frame resident in HBM (GpuFrameWarper::registerFrame). uint16 frames are kept as
    counts and sampled as count + offset; offset defaults to 1 for uint16 (the aligner's
    +1.0f, FrameWarpAligner.cpp:142) and 0 for float frames (already offset)."""

    def __init__(self, ref, frame, offset=None):
        self.ref = ref
        self.L = ref.L
        a = np.ascontiguousarray(frame)
        if a.dtype == np.uint16:
            dt = FA_U16
            off = 1.0 if offset is None else offset
        else:
            a = np.ascontiguousarray(a, dtype=np.float32)
            dt = FA_F32
            off = 0.0 if offset is None else offset
        if a.size != ref.n_px:
            raise FaError(-1, "Frame is incorrect size")
        h = C.c_void_p()
        _check(self.L.fa_register_frame(ref.h, _ptr(a), dt, float(off), C.byref(h)))
        self.h = h
        self.dtype = dt
        ref.ctx._adopt(self)

    def _release_handle(self):
        if getattr(self, "h", None):
            self.L.fa_deregister_frame(self.h)
            self.h = None

    release = _release_handle

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass

    def error(self, D):
        D = as_D(D, self.ref.nD, self.ref.n_dim)
        e = C.c_double()
        _check(self.L.fa_get_error(self.h, _ptr(D), C.byref(e)))
        return e.value

    def jacobian(self, D):
        D = as_D(D, self.ref.nD, self.ref.n_dim)
        jac = np.zeros(self.ref.nD * self.ref.n_dim)
        e = C.c_double()
        _check(self.L.fa_get_jacobian(self.h, _ptr(D), _ptr(jac), C.byref(e)))
        return e.value, jac

    def warp_interpolated(self, D, invalid_value=0.0):
        D = as_D(D, self.ref.nD, self.ref.n_dim)
        out = np.zeros(self.ref.n_px, np.float32)
        _check(self.L.fa_warp_interpolated(self.h, _ptr(D), _ptr(out), float(invalid_value)))
        return out.reshape(self.ref.shape)

    def warp_image(self, D, apply_offset=False):
        """Forward splat (warpImage): returns (wimg, coverage)."""
        D = as_D(D, self.ref.nD, self.ref.n_dim)
        w = np.zeros(self.ref.n_px, np.float32)
        c = np.zeros(self.ref.n_px, np.float32)
        _check(self.L.fa_warp_image(self.h, _ptr(D), _ptr(w), _ptr(c), 1 if apply_offset else 0))
        return w.reshape(self.ref.shape), c.reshape(self.ref.shape)

    def align(self, x0=None, **kw):
        n = self.ref.nD * self.ref.n_dim
        x = np.zeros(n) if x0 is None else np.array(x0, np.float64).reshape(n).copy()
        p = tr_params(**kw)
        s = AlignStats()
        _check(self.L.fa_align(self.h, C.byref(p), _ptr(x), C.byref(s)))
        return x, stats_dict(s)

    def align_noalloc(self, x0=None, **kw):
        """align() through fa_align_noalloc (no device allocation; the library's per-GPU
        scratch array)."""
        n = self.ref.nD * self.ref.n_dim
        x = np.zeros(n) if x0 is None else np.array(x0, np.float64).reshape(n).copy()
        p = tr_params(**kw)
        s = AlignStats()
        _check(self.L.fa_align_noalloc(self.h, C.byref(p), _ptr(x), C.byref(s)))
        return x, stats_dict(s)


class Batch:
    """Many frames aligned together: every trust-region round is one fused pass over all
    active frames plus one device update."""

    def __init__(self, frames, **kw):
        self.frames = list(frames)
        self.L = self.frames[0].L
        arr = (C.c_void_p * len(self.frames))(*[f.h.value for f in self.frames])
        self._arr = arr
        p = tr_params(**kw)
        h = C.c_void_p()
        _check(self.L.fa_batch_create(arr, len(self.frames), C.byref(p), C.byref(h)))
        self.h = h
        self.frames[0].ref.ctx._adopt(self)
        self.sizes = [f.ref.nD * f.ref.n_dim for f in self.frames]
        self.offsets = np.concatenate([[0], np.cumsum(self.sizes)]).astype(np.int64)

    def _release_handle(self):
        if getattr(self, "h", None):
            self.L.fa_batch_destroy(self.h)
            self.h = None

    close = _release_handle

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_timing(self, on=True):
        _check(self.L.fa_batch_set_timing(self.h, 1 if on else 0))
        self._timing = bool(on)

    def set_trace(self, max_records):
        """Records every trust-region evaluation of the following runs (TRACE_FIELDS, then
        x, g and p), at most max_records per frame; 0 turns tracing off."""
        _check(self.L.fa_batch_set_trace(self.h, int(max_records)))
        self._trace_cap = int(max_records)

    def trace(self):
        """Records of the last run: one (evaluations, TRACE_FIELDS + 3 n) array per frame."""
        cap = getattr(self, "_trace_cap", 0)
        nmax = max(self.sizes)
        rec = len(TRACE_FIELDS) + 3 * nmax
        out = np.zeros((len(self.frames), cap, rec))
        _check(self.L.fa_batch_get_trace(self.h, _ptr(out)))
        res = []
        for i, n in enumerate(self.sizes):
            k = int(np.count_nonzero(out[i, :, 0]))
            r = out[i, :k]
            res.append(np.concatenate([r[:, :len(TRACE_FIELDS)], r[:, len(TRACE_FIELDS):len(TRACE_FIELDS) + 3 * n]],
                                      axis=1))
        return res

    @property
    def tuning(self):
        t = Tuning()
        _check(self.L.fa_batch_get_tuning(self.h, C.byref(t)))
        return {f: getattr(t, f) for f, _ in Tuning._fields_}

    def set_tuning(self, **kw):
        t = Tuning(**self.tuning)
        for k, v in kw.items():
            setattr(t, k, int(v))
        _check(self.L.fa_batch_set_tuning(self.h, C.byref(t)))

    def run(self, x0=None):
        x = np.zeros(self.offsets[-1]) if x0 is None else np.concatenate(
            [np.asarray(v, np.float64).ravel() for v in x0])
        st = (AlignStats * len(self.frames))()
        tm = BatchTiming()
        _check(self.L.fa_batch_run(self.h, _ptr(x), st, C.byref(tm)))
        xs = [x[self.offsets[i]:self.offsets[i + 1]].copy() for i in range(len(self.frames))]
        timing = {f: getattr(tm, f) for f, _ in BatchTiming._fields_}
        return xs, [stats_dict(s) for s in st], timing

    def eval(self, Ds):
        D = np.concatenate([as_D(d, f.ref.nD, f.ref.n_dim).ravel() for d, f in zip(Ds, self.frames)])
        err = np.zeros(len(self.frames))
        jac = np.zeros(self.offsets[-1])
        _check(self.L.fa_batch_eval(self.h, _ptr(D), _ptr(err), _ptr(jac)))
        return err, [jac[self.offsets[i]:self.offsets[i + 1]].copy() for i in range(len(self.frames))]


class PinnedArray:
    """A numpy array in pinned host memory (fa_host_alloc), freed with the object."""

    def __init__(self, shape, dtype):
        self.L = load_library()
        dtype = np.dtype(dtype)
        n = int(np.prod(shape)) * dtype.itemsize
        p = C.c_void_p()
        _check(self.L.fa_host_alloc(n, C.byref(p)))
        self.p = p
        self.array = np.ctypeslib.as_array((C.c_uint8 * max(n, 1)).from_address(p.value))[:n] \
            .view(dtype).reshape(shape)

    def close(self):
        if getattr(self, "p", None):
            self.array = None
            self.L.fa_host_free(self.p)
            self.p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Continuous batching: frames submitted at any time join the trust-region rounds of the
    frames already in flight (a fixed number of slots) and leave when they finish."""

    def __init__(self, refs, slots, arena_bytes=0, **kw):
        refs = list(refs)
        self.L = refs[0].L
        self.ctx = refs[0].ctx
        arr = (C.c_void_p * len(refs))(*[r.h.value for r in refs])
        p = tr_params(**kw)
        h = C.c_void_p()
        _check(self.L.fa_server_create(self.ctx.h, arr, len(refs), int(slots), int(arena_bytes),
                                       C.byref(p), C.byref(h)))
        self.h = h
        self.slots = int(slots)
        self._live = {}  # tag -> (frame, x): kept alive until the result is taken
        self._buf = (ServerResult * max(self.slots, 64))()
        self.ctx._adopt(self)

    def _release_handle(self):
        if getattr(self, "h", None):
            self.L.fa_server_destroy(self.h)
            self.h = None
            self._live.clear()

    close = _release_handle

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, frame, tag, x0=None):
        n = frame.ref.nD * frame.ref.n_dim
        x = np.zeros(n) if x0 is None else np.array(x0, np.float64).reshape(n).copy()
        if tag in self._live:
            raise ValueError(f"tag {tag} already in flight")
        self._live[tag] = (frame, x)
        try:
            _check(self.L.fa_server_submit(self.h, frame.h, _ptr(x), int(tag)))
        except Exception:
            del self._live[tag]
            raise

    def submit_host(self, ref, data, tag, x0=None, offset=None):
        """A frame in host memory (ideally a PinnedArray), uploaded into the server's frame
        arena when a slot and arena room are free. data must stay unchanged until the result."""
        a = data if isinstance(data, np.ndarray) else np.asarray(data)
        if not a.flags.c_contiguous or a.dtype not in (np.uint16, np.float32):
            raise ValueError("host frames are C-contiguous uint16 or float32")
        if a.size != ref.n_px:
            raise FaError(-1, "Frame is incorrect size")
        dt = FA_U16 if a.dtype == np.uint16 else FA_F32
        off = (1.0 if dt == FA_U16 else 0.0) if offset is None else offset
        n = ref.nD * ref.n_dim
        x = np.zeros(n) if x0 is None else np.array(x0, np.float64).reshape(n).copy()
        if tag in self._live:
            raise ValueError(f"tag {tag} already in flight")
        self._live[tag] = (a, x)
        try:
            _check(self.L.fa_server_submit_host(self.h, ref.h, _ptr(a), dt, float(off), _ptr(x), int(tag)))
        except Exception:
            del self._live[tag]
            raise

    def memory(self):
        """(used, total) bytes of the frame arena."""
        u, t = C.c_uint64(), C.c_uint64()
        _check(self.L.fa_server_memory(self.h, C.byref(u), C.byref(t)))
        return u.value, t.value

    def step(self, rounds=1):
        """Collects the results of the lanes whose rounds finished, admits queued frames and
        launches `rounds` trust-region rounds on every lane with none in flight (waiting up to
        0.5 ms for a lane to finish when none had): returns (in_flight, queued,
        [(tag, x, stats), ...])."""
        a, q = C.c_int32(), C.c_int32()
        _check(self.L.fa_server_step(self.h, int(rounds), C.byref(a), C.byref(q)))
        out = []
        while True:
            n = C.c_int32()
            _check(self.L.fa_server_results(self.h, self._buf, len(self._buf), C.byref(n)))
            for i in range(n.value):
                r = self._buf[i]
                _, x = self._live.pop(r.tag)
                out.append((r.tag, x, stats_dict(r.stats)))
            if n.value < len(self._buf):
                break
        return a.value, q.value, out

    def run_stream(self, times, refs, frames, rounds=3, offsets=None):
        """The request loop in the library (fa_server_run_stream): frame i (a host array, ideally
        pinned, of refs[i]) arrives times[i] seconds after the call starts; returns (latency s
        per frame, [(x, stats)] per frame, counts). The server must hold no other frames."""
        n = len(times)
        t = np.ascontiguousarray(times, np.float64)
        hs = (C.c_void_p * n)(*[r.h for r in refs])
        arrs = []
        for a in frames:
            a = a if isinstance(a, np.ndarray) else np.asarray(a)
            if not a.flags.c_contiguous or a.dtype not in (np.uint16, np.float32):
                raise ValueError("host frames are C-contiguous uint16 or float32")
            arrs.append(a)
        for a, r in zip(arrs, refs):
            if a.size != r.n_px:
                raise FaError(-1, "Frame is incorrect size")
        ptrs = (C.c_void_p * n)(*[a.ctypes.data for a in arrs])
        dt = np.array([FA_U16 if a.dtype == np.uint16 else FA_F32 for a in arrs], np.int32)
        if offsets is None:
            off = np.array([1.0 if d == FA_U16 else 0.0 for d in dt], np.float32)
        else:
            off = np.ascontiguousarray(offsets, np.float32)
        sizes = np.array([r.nD * r.n_dim for r in refs], np.int64)
        xoff = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64) if n else np.zeros(0, np.int64)
        x = np.zeros(int(sizes.sum()) if n else 0)
        lat = np.zeros(n)
        st = (AlignStats * max(n, 1))()
        cnt = StreamCounts()
        _check(self.L.fa_server_run_stream(self.h, n, _ptr(t), hs, ptrs, _ptr(dt), _ptr(off), _ptr(x), _ptr(xoff),
                                           int(rounds), _ptr(lat), st, C.byref(cnt)))
        res = [(x[xoff[i]:xoff[i] + sizes[i]].copy(), stats_dict(st[i])) for i in range(n)]
        return lat, res, {f: getattr(cnt, f) for f, _ in StreamCounts._fields_}


class Aligner:
    """FrameWarpAligner on the device (FrameWarpAligner.cpp:46-223): raw reference and frames,
    optional blur, +1, rigid phase-correlation start, best of the starts, trust region."""

    def __init__(self, ctx, reference, type="warp", n_resampling_points=30, smoothing=0.0,
                 spatial_binning=1, line=1.0, interline=1.0, interframe=None):
        self.ctx = ctx
        self.L = ctx.L
        ref = np.ascontiguousarray(reference)
        if ref.ndim == 2:
            ref = ref[None]
        self.dtype = FA_U16 if ref.dtype == np.uint16 else FA_F32
        if self.dtype == FA_F32:
            ref = np.ascontiguousarray(ref, dtype=np.float32)
        nz, ny, nx = ref.shape
        self.shape = (nz, ny, nx)
        if interframe is None:
            interframe = ny * max(interline, line)
        sp = ScanParams(line, interline, interframe, nx, ny, nz, 0)
        rp = RealignParams(REALIGN_TYPES[type] if isinstance(type, str) else int(type), spatial_binning,
                           n_resampling_points, 0, smoothing)
        h = C.c_void_p()
        _check(self.L.fa_aligner_create(ctx.h, C.byref(rp), C.byref(sp), _ptr(ref), self.dtype, C.byref(h)))
        self.h = h
        ctx._adopt(self)
        nD = C.c_int32()
        _check(self.L.fa_aligner_reference(self.h, None, C.byref(nD)))
        self.nD = nD.value
        self.n_dim = 3 if nz > 1 else 2

    def _release_handle(self):
        if getattr(self, "h", None):
            self.L.fa_aligner_destroy(self.h)
            self.h = None

    close = _release_handle

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _raw(self, frame):
        a = np.ascontiguousarray(frame)
        a = a if a.dtype == np.uint16 else np.ascontiguousarray(a, dtype=np.float32)
        if a.size != int(np.prod(self.shape)):
            raise FaError(-1, "Frame is incorrect size")
        return a, (FA_U16 if a.dtype == np.uint16 else FA_F32)

    def preprocess(self, frame):
        """smoothStack(frame) + 1, as the warper receives it."""
        a, dt = self._raw(frame)
        out = np.zeros(self.shape, np.float32)
        _check(self.L.fa_aligner_preprocess(self.h, _ptr(a), dt, _ptr(out)))
        return out

    def rigid_shift(self, frame):
        a, dt = self._raw(frame)
        o = np.zeros(3)
        _check(self.L.fa_aligner_rigid_shift(self.h, _ptr(a), dt, _ptr(o)))
        return tuple(float(v) for v in o)

    def add_frames(self, frames, outputs=(), **kw):
        """addFrame for each raw frame (one dtype): returns (D list (nD x 3), result dicts, dict
        of the requested output images: any of 'realigned', 'preserving' (uint16), 'mask',
        'warped' (float), each n x nz x ny x nx)."""
        arrs = [self._raw(f) for f in frames]
        dts = {dt for _, dt in arrs}
        if len(dts) != 1:
            raise FaError(-1, "frames of one dtype expected")
        n = len(arrs)
        ptrs = (C.c_void_p * n)(*[a.ctypes.data for a, _ in arrs])
        D = np.zeros((n, self.nD, 3))
        res = (AddResult * n)()
        imgs = {}
        for k in outputs:
            imgs[k] = np.zeros((n,) + self.shape, np.uint16 if k in ("realigned", "preserving") else np.float32)
        o = AddOutputs(*[imgs[k].ctypes.data if k in imgs else None
                         for k in ("realigned", "preserving", "mask", "warped")])
        p = tr_params(**kw)
        _check(self.L.fa_aligner_add_frames(self.h, ptrs, n, dts.pop(), C.byref(p), _ptr(D), C.byref(o), res))
        results = [dict(start=r.start, rigid_shift=tuple(r.rigid_shift), start_error=tuple(r.start_error),
                        stats=stats_dict(r.stats), correlation=r.correlation,
                        unaligned_correlation=r.unaligned_correlation, coverage=r.coverage) for r in res]
        return [D[i] for i in range(n)], results, imgs

    def clear_temp(self):
        """FrameWarpAligner::clearTemp: frees what add_frames keeps between calls and the
        warper's cached working spaces."""
        _check(self.L.fa_aligner_clear_temp(self.h))

    def realign_as_frame(self, D, frame, interpolate_missing=False, cv_type=0):
        """FrameWarpAligner::realignAsFrame: frame (as given, no blur or offset) realigned with D
        (nD x 3). interpolate_missing: the interpolated warp (float32, invalid samples 0); else
        the splat's weighted values converted to cv_type (0 = CV_8U, 2 = CV_16U, 5 = CV_32F)."""
        a, dt = self._raw(frame)
        if interpolate_missing:
            out_dtype = np.float32
        else:
            out_dtype = {0: np.uint8, 2: np.uint16, 5: np.float32}.get(int(cv_type))
            if out_dtype is None:
                raise FaError(-1, "cv_type must be 0 (CV_8U), 2 (CV_16U) or 5 (CV_32F)")
        out = np.zeros(self.shape, out_dtype)
        _check(self.L.fa_aligner_realign_as_frame(self.h, _ptr(as_D(D, self.nD, self.n_dim)), _ptr(a), dt,
                                                  int(bool(interpolate_missing)), int(cv_type), _ptr(out)))
        return out

    def shift_pixel(self, D, x, y, z):
        """FrameWarpAligner::shiftPixel for a frame aligned to D."""
        ref = C.c_void_p()
        _check(self.L.fa_aligner_reference(self.h, C.byref(ref), None))
        p = np.array([x, y, z], np.float64)
        _check(self.L.fa_shift_pixel(ref, _ptr(as_D(D, self.nD, self.n_dim)), _ptr(p)))
        return tuple(float(v) for v in p)


def write_realignment_info(filename, D, unaligned_correlation, correlation, coverage, n_x, n_y, n_z=1,
                           line=1.0, interline=1.0, interframe=None):
    """FrameWarpAligner::writeRealignmentInfo: the version-2 CSV of the scan timing and, per
    frame, the correlations, coverage and node displacements (D: n x nD x 3)."""
    L = load_library()
    D = np.ascontiguousarray(D, np.float64)
    n, nD = D.shape[0], D.shape[1]
    if interframe is None:
        interframe = n_y * max(interline, line)
    sp = ScanParams(line, interline, interframe, n_x, n_y, n_z, 0)
    arr = [np.ascontiguousarray(v, np.float64) for v in (unaligned_correlation, correlation, coverage)]
    _check(L.fa_write_realignment_info(os.fsencode(filename), C.byref(sp), n, nD, _ptr(D), *[_ptr(a) for a in arr]))


class HipFrameWarper:
    """The reference's AbstractFrameWarper interface (FrameWarper.h:95-143) on the GPU.
    Frames are keyed by their data pointer, as the reference keys them by cv::Mat::data
    (GpuFrameWarper.h:55)."""

    def __init__(self, device=0, ctx=None):
        self.ctx = ctx or Context(device)
        self.ref = None
        self.frames = {}

    def setReference(self, reference, nD, line_duration=1.0, interline_duration=1.0,
                     interframe_duration=None):
        self.ref = Reference(self.ctx, reference, nD, line_duration, interline_duration,
                             interframe_duration)
        self.n_dim = self.ref.n_dim
        self.dims = list(self.ref.shape)
        self.nD = nD

    @property
    def H(self):
        return self.ref.H

    def _key(self, frame):
        return frame.__array_interface__["data"][0]

    def registerFrame(self, frame):
        self.frames[self._key(frame)] = self.ref.register(frame, offset=0.0)

    def deregisterFrame(self, frame):
        f = self.frames.pop(self._key(frame), None)
        if f is not None:
            f.release()

    def _frame(self, frame):
        f = self.frames.get(self._key(frame))
        if f is None:
            raise FaError(-3, "Unrecognised frame")
        return f

    def getError(self, frame, D):
        return self._frame(frame).error(D)

    def getJacobian(self, frame, D):
        return self._frame(frame).jacobian(D)[1]

    def warpImageInterpolated(self, frame, D, invalid_value=0):
        return self._frame(frame).warp_interpolated(D, invalid_value)

    def warpImage(self, frame, D):
        return self._frame(frame).warp_image(D)

    def warpPoint(self, D, x, y, z, spatial_binning=1):
        return self.ref.warp_point(D, x, y, z, spatial_binning)

    def clearTemp(self):
        if self.ref is not None:
            self.ref.clear_temp()
