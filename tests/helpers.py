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


# Parity cases shared by the GPU tests: (name, shape, nD, scan, generator options). Each is
# well posed: the oracle's converged displacement moves by far less than the 0.02 px
# tolerance when the start is perturbed (tests/test_parity_cases.py), so comparing two
# implementations' converged displacements is meaningful.
PARITY_CASES = [
    ("2d", (1, 96, 80), 9, dict(line=1.0, interline=1.0), {}),
    ("2d_interline", (1, 64, 72), 13, dict(line=1.0, interline=1.3), {}),
    ("3d", (6, 40, 36), 6 * 3 + 1, dict(line=1.0, interline=1.0), dict(n_blobs=120, sigma=(1.5, 4.0))),
]
