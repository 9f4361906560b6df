"""Command front end (fa_amd.commands, the reference's FrameAlignerMex.cpp:18-162): registry and
argument handling, which run before any device work."""
import numpy as np
import pytest

import fa_amd as F
from fa_amd.commands import frame_aligner


def test_registry_and_errors():
    a = frame_aligner(dict(type=2, n_resampling_points=4), dict(n_x=32, n_y=16), 10)
    b = frame_aligner(dict(), dict(n_x=32, n_y=16), 10)
    assert b == a + 1
    frame_aligner(a, "Delete")
    c = frame_aligner(dict(), dict(n_x=8, n_y=8), 1)
    assert c == a                       # the first empty place is reused
    with pytest.raises(F.FaError, match="Invalid aligner index"):
        frame_aligner(a + 100, "NumTasksRemaining")
    with pytest.raises(F.FaError, match="command string"):
        frame_aligner(b, 3)
    with pytest.raises(F.FaError, match="Reference imaging not yet set"):
        frame_aligner(b, "AddFrame", 0, np.zeros((16, 32), np.uint16))
    with pytest.raises(F.FaError, match="Result not ready"):
        frame_aligner(b, "GetRealignedFrame", 0)
    assert frame_aligner(b, "NumTasksRemaining") == 0
    with pytest.raises(F.FaError):
        frame_aligner(dict(spatial_binning=0), dict(n_x=8, n_y=8), 1)
    with pytest.raises(F.FaError):
        frame_aligner(b, "SetNumFrames", 0)
    for i in (b, c):
        frame_aligner(i, "Delete")
