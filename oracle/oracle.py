"""ctypes front end of the CPU oracle (oracle/fa_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker. The product never imports it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "build", "libfa_oracle.so")
# the same source built with the reference's Release flags; times the CPU baseline only
FAST_LIB_PATH = os.path.join(HERE, "build", "libfa_oracle_fast.so")


def build(force=False):
    """Compiles the oracle with its Makefile (gcc, no GPU needed)."""
    if force or not (os.path.exists(LIB_PATH) and os.path.exists(FAST_LIB_PATH)):
        subprocess.run(["make", "-s", "-C", HERE, "-B" if force else "all"], check=True)
    return LIB_PATH


class TrParams(C.Structure):
    _fields_ = [("radius", C.c_double), ("min_delta", C.c_double), ("rel_delta", C.c_double),
                ("max_iter", C.c_int), ("max_evals", C.c_int), ("sub_eps", C.c_double),
                ("sub_max_iter", C.c_int), ("skip_repeat", C.c_int), ("dense", C.c_int)]


class TrStats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("evaluations", C.c_int), ("rejected", C.c_int),
                ("status", C.c_int), ("f_final", C.c_double), ("radius_final", C.c_double)]


VALUE_FN = C.CFUNCTYPE(C.c_double, C.c_void_p, C.POINTER(C.c_double))
GRADIENT_FN = C.CFUNCTYPE(None, C.c_void_p, C.POINTER(C.c_double))


class TrModel(C.Structure):
    _fields_ = [("nb", C.c_int), ("m", C.c_int), ("diag", C.c_void_p), ("sub", C.c_void_p),
                ("value", VALUE_FN), ("gradient", GRADIENT_FN), ("user", C.c_void_p)]


# Per-evaluation trace record of the trust region (fa_oracle.h ORC_T_*), then x, g, p.
TRACE_FIELDS = ["evals", "f_new", "f", "pred", "meas", "rho", "accept", "radius", "iterations",
                "status", "lambda", "sub_iters", "pnorm", "gnorm", "gp", "rejected"]
N_TRACE = len(TRACE_FIELDS)

_libs = {}


def lib(fast=False):
    """The checker library, or with fast=True the Release-flag build used for timing."""
    path = FAST_LIB_PATH if fast else LIB_PATH
    if path not in _libs:
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        P = C.c_void_p
        dp = C.POINTER(C.c_double)
        L.orc_create.restype = P
        L.orc_create.argtypes = [C.c_int]
        L.orc_destroy.argtypes = [P]
        L.orc_last_error.restype = C.c_char_p
        L.orc_set_reference.argtypes = [P, P, C.c_int, C.c_int, C.c_int, C.c_int,
                                        C.c_double, C.c_double, C.c_double]
        L.orc_n_dim.argtypes = [P]
        L.orc_get_scan.argtypes = [P, P]
        L.orc_get_node_map.argtypes = [P, P, P, P]
        L.orc_get_gradient.argtypes = [P, P, P, P]
        L.orc_get_vi_ranges.argtypes = [P, P]
        L.orc_get_hessian.argtypes = [P, P]
        L.orc_get_error.restype = C.c_double
        L.orc_get_error.argtypes = [P, P, P, P, P]
        L.orc_get_jacobian.restype = C.c_double
        L.orc_get_jacobian.argtypes = [P, P, P, P]
        L.orc_warp_interpolated.argtypes = [P, P, P, C.c_double, P]
        L.orc_warp_image.argtypes = [P, P, P, P, P]
        L.orc_warp_point.argtypes = [P, P, C.c_int, C.c_int, C.c_int, C.c_int, P]
        L.orc_tr_default.argtypes = [C.POINTER(TrParams)]
        L.orc_align.argtypes = [P, P, C.POINTER(TrParams), P, C.POINTER(TrStats)]
        L.orc_align_trace.argtypes = [P, P, C.POINTER(TrParams), P, C.POINTER(TrStats), P, C.c_int]
        L.orc_tr_run.argtypes = [C.POINTER(TrModel), C.POINTER(TrParams), P, C.POINTER(TrStats), P, C.c_int]
        L.orc_align_many.argtypes = [P, P, C.c_int, C.POINTER(TrParams), P, C.POINTER(TrStats),
                                     P, C.c_int]
        L.orc_bt_cholesky.argtypes = [C.c_int, C.c_int, P, P, C.c_double, P, P]
        L.orc_bt_solve.argtypes = [C.c_int, C.c_int, P, P, P, P]
        L.orc_tr_subproblem_bt.argtypes = [C.c_int, C.c_int, P, P, P, C.c_double, C.c_double,
                                           C.c_int, P, dp]
        L.orc_tr_subproblem_dense.argtypes = [C.c_int, P, P, C.c_double, C.c_double, C.c_int,
                                              P, dp]
        _libs[path] = L
    return _libs[path]


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def tr_params(**kw):
    p = TrParams()
    lib().orc_tr_default(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def stats_dict(s):
    return {f: getattr(s, f) for f, _ in TrStats._fields_}


class Oracle:
    """CpuFrameWarper restated (FrameWarper.h:95-159). fp64=False is the
    reference-faithful float-storage mode, fp64=True the all-double mode."""

    def __init__(self, fp64=False, fast=False):
        self.L = lib(fast)
        self.h = self.L.orc_create(1 if fp64 else 0)
        self.fp64 = fp64

    def __del__(self):
        try:
            if self.h:
                self.L.orc_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def set_reference(self, ref, nD, line=1.0, interline=1.0, interframe=None):
        ref = np.ascontiguousarray(ref, dtype=np.float32)
        if ref.ndim == 2:
            ref = ref[None]
        nz, ny, nx = ref.shape
        if interframe is None:
            interframe = ny * max(interline, line)
        self.ref = ref
        rc = self.L.orc_set_reference(self.h, _ptr(ref), nz, ny, nx, nD, line, interline, interframe)
        if rc != 0:
            raise RuntimeError(self.L.orc_last_error().decode())
        self.shape = ref.shape
        self.nD = nD
        self.n_dim = self.L.orc_n_dim(self.h)
        self.n_px = nz * ny * nx
        return self

    def scan(self):
        o = np.zeros(6)
        self.L.orc_get_scan(self.h, _ptr(o))
        return dict(zip(["line", "interline", "pixel", "frame", "interframe", "stack"], o))

    def node_map(self):
        Di = np.zeros(self.n_px, np.uint16)
        Df = np.zeros(self.n_px, np.float64)
        Dr = np.zeros((self.nD, 2), np.int32)
        self.L.orc_get_node_map(self.h, _ptr(Di), _ptr(Df), _ptr(Dr))
        return Di, Df, Dr

    def gradient(self):
        g = [np.zeros(self.n_px) for _ in range(3)]
        self.L.orc_get_gradient(self.h, *[_ptr(a) for a in g])
        return [a.reshape(self.shape) for a in g]

    def vi_ranges(self):
        r = np.zeros((self.nD, 2), np.int64)
        self.L.orc_get_vi_ranges(self.h, _ptr(r))
        return r

    def hessian(self):
        n = self.nD * self.n_dim
        H = np.zeros((n, n))
        self.L.orc_get_hessian(self.h, _ptr(H))
        return H

    def _D(self, D):
        D = np.ascontiguousarray(D, dtype=np.float64)
        if D.ndim == 1:
            D = self.col2D(D)
        assert D.shape == (self.nD, 3)
        return D

    def col2D(self, x):
        x = np.asarray(x, np.float64).reshape(self.nD, self.n_dim)
        D = np.zeros((self.nD, 3))
        D[:, :self.n_dim] = x
        return D

    def _frame(self, frame):
        f = np.ascontiguousarray(frame, dtype=np.float32).reshape(self.shape)
        return f

    def get_error(self, frame, D, images=False):
        f = self._frame(frame)
        D = self._D(D)
        if images:
            err = np.zeros(self.n_px)
            wimg = np.zeros(self.n_px)
            e = self.L.orc_get_error(self.h, _ptr(f), _ptr(D), _ptr(err), _ptr(wimg))
            return e, err.reshape(self.shape), wimg.reshape(self.shape)
        return self.L.orc_get_error(self.h, _ptr(f), _ptr(D), None, None)

    def get_jacobian(self, frame, D):
        f = self._frame(frame)
        D = self._D(D)
        jac = np.zeros(self.nD * self.n_dim)
        e = self.L.orc_get_jacobian(self.h, _ptr(f), _ptr(D), _ptr(jac))
        return e, jac

    def warp_interpolated(self, frame, D, invalid_value=0.0):
        f = self._frame(frame)
        D = self._D(D)
        out = np.zeros(self.n_px)
        self.L.orc_warp_interpolated(self.h, _ptr(f), _ptr(D), float(invalid_value), _ptr(out))
        return out.reshape(self.shape)

    def warp_image(self, frame, D):
        f = self._frame(frame)
        D = self._D(D)
        w = np.zeros(self.n_px)
        c = np.zeros(self.n_px)
        self.L.orc_warp_image(self.h, _ptr(f), _ptr(D), _ptr(w), _ptr(c))
        return w.reshape(self.shape), c.reshape(self.shape)

    def warp_point(self, D, x, y, z, binning=1):
        D = self._D(D)
        o = np.zeros(3)
        self.L.orc_warp_point(self.h, _ptr(D), x, y, z, binning, _ptr(o))
        return o

    def align(self, frame, x0=None, **kw):
        f = self._frame(frame)
        n = self.nD * self.n_dim
        x = np.zeros(n) if x0 is None else np.array(x0, np.float64).copy()
        p = tr_params(**kw)
        s = TrStats()
        self.L.orc_align(self.h, _ptr(f), C.byref(p), _ptr(x), C.byref(s))
        return x, stats_dict(s)

    def align_trace(self, frame, cap=64, x0=None, **kw):
        """align() with the per-evaluation trace: (x, stats, records), records of shape
        (evaluations, N_TRACE + 3n) (fields TRACE_FIELDS, then x, g, p after each record)."""
        f = self._frame(frame)
        n = self.nD * self.n_dim
        x = np.zeros(n) if x0 is None else np.array(x0, np.float64).copy()
        p = tr_params(**kw)
        s = TrStats()
        tr = np.zeros((cap, N_TRACE + 3 * n))
        self.L.orc_align_trace(self.h, _ptr(f), C.byref(p), _ptr(x), C.byref(s), _ptr(tr), cap)
        return x, stats_dict(s), tr[:min(cap, s.evaluations)]

    def align_many(self, frames, x0=None, threads=None, **kw):
        frames = np.ascontiguousarray(frames, dtype=np.float32)
        nf = frames.shape[0]
        assert frames.size == nf * self.n_px
        n = self.nD * self.n_dim
        x = np.zeros((nf, n)) if x0 is None else np.array(x0, np.float64).copy()
        p = tr_params(**kw)
        st = (TrStats * nf)()
        secs = np.zeros(nf)
        threads = threads or os.cpu_count() or 1
        self.L.orc_align_many(self.h, _ptr(frames), nf, C.byref(p), _ptr(x), st, _ptr(secs), threads)
        return x, [stats_dict(s) for s in st], secs


def bt_cholesky(diag, sub, lam=0.0):
    nb, m, _ = diag.shape
    Ld = np.zeros_like(diag)
    Ls = np.zeros_like(diag)
    ok = lib().orc_bt_cholesky(nb, m, _ptr(np.ascontiguousarray(diag)),
                               _ptr(np.ascontiguousarray(sub)), lam, _ptr(Ld), _ptr(Ls))
    return bool(ok), Ld, Ls


def bt_solve(Ld, Ls, b):
    nb, m, _ = Ld.shape
    x = np.zeros(nb * m)
    lib().orc_bt_solve(nb, m, _ptr(Ld), _ptr(Ls), _ptr(np.ascontiguousarray(b, np.float64)), _ptr(x))
    return x


def tr_subproblem_bt(diag, sub, g, radius, eps=0.1, max_iter=20):
    nb, m, _ = diag.shape
    p = np.zeros(nb * m)
    lam = C.c_double()
    it = lib().orc_tr_subproblem_bt(nb, m, _ptr(np.ascontiguousarray(diag)),
                                    _ptr(np.ascontiguousarray(sub)),
                                    _ptr(np.ascontiguousarray(g, np.float64)), radius, eps,
                                    max_iter, _ptr(p), C.byref(lam))
    return p, lam.value, it


def tr_subproblem_dense(B, g, radius, eps=0.1, max_iter=20):
    n = B.shape[0]
    p = np.zeros(n)
    lam = C.c_double()
    it = lib().orc_tr_subproblem_dense(n, _ptr(np.ascontiguousarray(B, np.float64)),
                                       _ptr(np.ascontiguousarray(g, np.float64)), radius, eps,
                                       max_iter, _ptr(p), C.byref(lam))
    return p, lam.value, it


def tr_run(diag, sub, value, gradient, x0, cap=64, **kw):
    """The oracle's trust-region loop over a closed-form model: value(x) -> f, gradient(x) -> g
    (numpy), constant block-tridiagonal B (diag: nb x m x m, sub: nb-1 x m x m). Returns
    (x, stats, records) as Oracle.align_trace."""
    diag = np.ascontiguousarray(diag, np.float64)
    nb, m, _ = diag.shape
    sub = np.ascontiguousarray(sub, np.float64).reshape(max(nb - 1, 0), m, m)
    subp = sub if sub.size else np.zeros((1, m, m))
    n = nb * m
    last = {}

    def _value(_u, xp):
        xv = np.ctypeslib.as_array(xp, shape=(n,)).copy()
        last["x"] = xv
        return float(value(xv))

    def _gradient(_u, gp):
        np.ctypeslib.as_array(gp, shape=(n,))[:] = gradient(last["x"])

    vf, gf = VALUE_FN(_value), GRADIENT_FN(_gradient)
    M = TrModel(nb, m, _ptr(diag), _ptr(subp), vf, gf, None)
    x = np.array(x0, np.float64).reshape(n).copy()
    p = tr_params(**kw)
    s = TrStats()
    tr = np.zeros((cap, N_TRACE + 3 * n))
    lib().orc_tr_run(C.byref(M), C.byref(p), _ptr(x), C.byref(s), _ptr(tr), cap)
    return x, stats_dict(s), tr[:min(cap, s.evaluations)]
