"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py from the oracle):
the oracle must reproduce them exactly (regression), and the HIP path must match them to the
parity tolerances."""
import glob
import os

import numpy as np
import pytest

import oracle as O
from helpers import rel_err

HERE = os.path.dirname(os.path.abspath(__file__))
FILES = sorted(glob.glob(os.path.join(HERE, "golden", "*.npz")))


def load(path):
    return dict(np.load(path, allow_pickle=False))


@pytest.mark.parametrize("path", FILES, ids=[os.path.basename(f) for f in FILES])
def test_oracle_reproduces_golden(path):
    g = load(path)
    nD = int(g["nD"])
    o = O.Oracle().set_reference(g["reference"], nD, line=1.0, interline=float(g["interline"]))
    Di, Df, Dr = o.node_map()
    assert np.array_equal(Di, g["Di"]) and np.array_equal(Df.astype(np.float32), g["Df"])
    assert np.array_equal(Dr, g["D_range"])
    gx, gy, gz = o.gradient()
    assert np.array_equal(gx.astype(np.float32), g["gx"]) and np.array_equal(gy.astype(np.float32), g["gy"])
    assert np.array_equal(o.hessian(), g["H"])
    err, jac = o.get_jacobian(g["frame"], g["D"])
    assert err == g["err"] and np.array_equal(jac, g["jac"])
    x, _ = o.align(g["frame"], rel_delta=1e-4, max_iter=50, skip_repeat=1)
    assert np.array_equal(x, g["x"])


@pytest.mark.gpu
@pytest.mark.parametrize("path", FILES, ids=[os.path.basename(f) for f in FILES])
def test_hip_matches_golden(path):
    import fa_amd as F
    g = load(path)
    nD = int(g["nD"])
    ctx = F.Context(0)
    ref = F.Reference(ctx, g["reference"], nD, line=1.0, interline=float(g["interline"]))
    Di, Df, Dr = ref.node_map()
    assert np.array_equal(Di, g["Di"]) and np.array_equal(Df, g["Df"]) and np.array_equal(Dr, g["D_range"])
    gx, gy, gz = ref.gradient()
    assert np.array_equal(gx, g["gx"]) and np.array_equal(gy, g["gy"])
    assert rel_err(ref.H, g["H"]) < 1e-3
    fr = ref.register(g["frame"])
    err, jac = fr.jacobian(g["D"])
    assert abs(err - g["err"]) <= 1e-4 * g["err"]
    assert rel_err(jac, g["jac"]) < 1e-3
    w = fr.warp_interpolated(g["D"], -1.0)
    assert np.array_equal(w == -1.0, g["warp"] == -1.0)
    x, _ = fr.align(rel_delta=1e-4, max_iter=50, skip_repeat=1)
    assert np.max(np.abs(x - g["x"])) <= 0.02
    ctx.close()
