"""The C++ mirror of the reference's warper interface (frame-aligner_amd/host/HipFrameWarper.hpp),
driven as FrameWarpAligner::addFrame drives a warper, against the oracle."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from helpers import rel_err, small_stack

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "frame-aligner_amd", "lib", "fa_host_check")


# well-posed cases: the converged displacement is stable under start perturbations
@pytest.mark.parametrize("shape,nD,gen", [((1, 80, 72), 9, {}),
                                          ((4, 32, 28), 4 * 3 + 1, dict(n_blobs=120, sigma=(1.5, 4.0)))])
def test_cpp_host_mirror(tmp_path, shape, nD, gen):
    assert os.path.exists(EXE), "build frame-aligner_amd/build.sh first"
    st = small_stack(shape, nD=nD, n_frames=1, seed=31, **gen)
    ref, fr = st["ref_w"], st["frames_w"][0]
    D = st["D_true"][0].copy()
    if shape[0] == 1:
        D[:, 2] = 0
    (tmp_path / "meta.txt").write_text(f"{shape[0]} {shape[1]} {shape[2]} {nD}\n")
    ref.astype(np.float32).tofile(tmp_path / "ref.f32")
    fr.astype(np.float32).tofile(tmp_path / "frame.f32")
    D.astype(np.float64).tofile(tmp_path / "D.f64")
    r = subprocess.run([EXE, str(tmp_path)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    o = O.Oracle().set_reference(ref, nD)
    e_o, j_o = o.get_jacobian(fr, D)
    err = np.fromfile(tmp_path / "err.f64")[0]
    jac = np.fromfile(tmp_path / "jac.f64")
    assert abs(err - e_o) <= 1e-4 * e_o
    assert rel_err(jac, j_o) < 1e-3
    x = np.fromfile(tmp_path / "x.f64")
    x_o, _ = o.align(fr, rel_delta=1e-4, max_iter=50, skip_repeat=1)
    assert np.max(np.abs(x - x_o)) <= 0.02
    w = np.fromfile(tmp_path / "warp.f32", np.float32).reshape(shape)
    w_o = o.warp_interpolated(fr, o.col2D(x), 0.0)
    assert np.array_equal(w == 0, w_o == 0)
    assert np.max(np.abs(w - w_o)) <= 1e-4 * np.max(np.abs(w_o))
    evals, status, unknown = map(int, (tmp_path / "stats.txt").read_text().split())
    assert evals >= 2 and 1 <= status <= 6 and unknown == 1
