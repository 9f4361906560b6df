"""Writes the golden fixtures of tests/golden/ from the CPU oracle.

The reference ships no golden vectors (it has no tests), so these are regression fixtures:
inputs and the oracle's outputs on small seeded cases (reference-faithful mode), kept so
that the oracle and the HIP path can both be checked against fixed numbers without
regenerating anything. Run from the repository root: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path[:0] = [os.path.join(ROOT, "oracle"), os.path.join(ROOT, "frame-aligner_amd")]
import oracle as O  # noqa: E402
from fa_amd import synth  # noqa: E402

CASES = {
    "g2d": dict(shape=(1, 40, 36), nD=7, seed=101, n_blobs=20, interline=1.0),
    "g2d_interline": dict(shape=(1, 33, 29), nD=9, seed=102, n_blobs=15, interline=1.2),
    "g3d": dict(shape=(4, 18, 16), nD=4 * 2 + 1, seed=103, n_blobs=12, interline=1.0),
}


def make(name, shape, nD, seed, n_blobs, interline):
    st = synth.make_stack(shape, 1, nD, seed=seed, n_blobs=n_blobs)
    ref = synth.with_offset(st["reference"])
    frame = synth.with_offset(st["frames"][0])
    D = st["D_true"][0].copy()
    if shape[0] == 1:
        D[:, 2] = 0
    o = O.Oracle().set_reference(ref, nD, line=1.0, interline=interline)
    Di, Df, Dr = o.node_map()
    gx, gy, gz = o.gradient()
    err, jac = o.get_jacobian(frame, D)
    warp = o.warp_interpolated(frame, D, -1.0)
    x, stats = o.align(frame, rel_delta=1e-4, max_iter=50, skip_repeat=1)
    np.savez_compressed(
        os.path.join(os.path.dirname(os.path.abspath(__file__)), f"{name}.npz"),
        reference=ref, frame=frame, D=D, nD=nD, interline=interline,
        Di=Di, Df=Df.astype(np.float32), D_range=Dr,
        gx=gx.astype(np.float32), gy=gy.astype(np.float32), gz=gz.astype(np.float32),
        H=o.hessian(), err=err, jac=jac, warp=warp.astype(np.float32), x=x,
        evaluations=stats["evaluations"])


if __name__ == "__main__":
    for k, v in CASES.items():
        make(k, **v)
        print("wrote", k)
