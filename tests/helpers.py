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


# Cases with many nodes: more node blocks than the 64 lanes of a trust-region wave take in one
# level of the cyclic-reduction factorisation (nD = 150 in 2-D, 161 in 3-D). The sparse blob
# phantom leaves some node intervals nearly featureless, so these cases start the trust region
# at the generator's displacement, from where they are well posed (tests/test_parity_cases.py).
MANY_NODE_CASES = [
    ("2d_many_nodes", (1, 160, 64), 150, dict(line=1.0, interline=1.0),
     dict(n_blobs=80, walk_sigma=0.05, sin_amp=1.5)),
    ("3d_many_nodes", (4, 64, 64), 4 * 40 + 1, dict(line=1.0, interline=1.0),
     dict(n_blobs=200, sigma=(1.5, 4.0), walk_sigma=0.05, sin_amp=1.5, z_amp=0.5)),
]

# A chain of 1539 parameters (3-D, 513 nodes: 64 per plane), long enough for the batched
# eigen-space subproblem kernels.
LONG_CHAIN_CASE = ("3d_long_chain", (8, 64, 64), 8 * 64 + 1, dict(line=1.0, interline=1.0),
                   dict(n_blobs=300, sigma=(1.5, 4.0), walk_sigma=0.05, sin_amp=1.5, z_amp=0.5))


def truth_start(st, f, n_dim):
    """Parameter vector (D2col layout) of frame f's generator displacement."""
    return st["D_true"][f][:, :n_dim].reshape(-1).copy()
