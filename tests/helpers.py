"""Shared fixtures for the parity tests: seeded stacks and tolerance checks."""
import numpy as np

from fa_amd import synth


def small_stack(shape=(1, 96, 80), n_frames=3, nD=9, seed=7, dtype="float32", n_blobs=40, **kw):
    st = synth.make_stack(shape, n_frames, nD, seed=seed, n_blobs=n_blobs, dtype=dtype, **kw)
    st["ref_w"] = synth.with_offset(st["reference"])
    st["frames_w"] = np.stack([synth.with_offset(f) for f in st["frames"]])
    return st


def rel_err(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
