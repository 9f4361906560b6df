"""The command front end end to end on the GPU: frames added asynchronously are aligned in
batches, and each result equals the aligner's own result for that frame."""
import time

import numpy as np
import pytest

import fa_amd as F
from fa_amd import synth
from fa_amd.commands import frame_aligner

pytestmark = pytest.mark.gpu


def test_commands_end_to_end(tmp_path):
    shape, n_res = (1, 64, 72), 6
    st = synth.make_stack(shape, 6, n_res + 1, seed=14, n_blobs=70, dtype="uint16")
    idx = frame_aligner(dict(type=2, n_resampling_points=n_res), dict(n_x=72, n_y=64), 6)
    frame_aligner(idx, "SetReference", 0, st["reference"][0])
    for t in range(6):
        frame_aligner(idx, "AddFrame", t, st["frames"][t][0])
        if t == 2:
            time.sleep(0.05)  # frames arriving while a batch runs join the next batch
    t0 = time.time()
    while frame_aligner(idx, "NumTasksRemaining") > 0:
        assert time.time() - t0 < 60
        time.sleep(0.01)
    ctx = F.Context(0)
    al = F.Aligner(ctx, st["reference"], n_resampling_points=n_res)
    D, res, imgs = al.add_frames(list(st["frames"]), outputs=("realigned",))
    for t in range(6):
        img = frame_aligner(idx, "GetRealignedFrame", t)
        assert img.shape == (64, 72) and img.dtype == np.uint16
        assert np.array_equal(img, imgs["realigned"][t][0])
        r = frame_aligner(idx, "GetRealignmentResult", t)
        assert np.array_equal(r["D"], D[t]) and r["correlation"] == res[t]["correlation"]
    path = tmp_path / "info.csv"
    frame_aligner(idx, "WriteRealignmentInfo", str(path))
    lines = path.read_text().splitlines()
    assert len(lines) == 8 + 6 and lines[8].startswith("0,")
    frame_aligner(idx, "Delete")
    with pytest.raises(F.FaError):
        frame_aligner(idx, "NumTasksRemaining")
    al.close()
    ctx.close()
