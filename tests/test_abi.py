"""The C ABI library loads, exports every entry point include/frame_aligner.h declares,
and reports errors instead of crashing when no GPU is present (no compute calls here)."""
import ctypes as C
import os
import re

import pytest

import fa_amd as F

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "frame_aligner.h")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fa_[a-z_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared()
    assert "fa_set_reference" in names and "fa_batch_run" in names
    assert len(names) >= 20


def test_library_exports_every_declared_symbol():
    L = F.load_library()
    for name in declared():
        assert hasattr(L, name), name
    assert set(declared()) == set(F.EXPORTS)


def test_defaults_match_reference_call():
    p = F.tr_params()
    assert p.radius == 40.0 and p.min_delta == 1e-8  # FrameWarpAligner.cpp:205-209
    assert p.max_iter == 0 and p.sub_eps == 0.1 and p.sub_max_iter == 20


def _gpu_visible():
    return os.path.exists("/dev/kfd")


@pytest.mark.skipif(_gpu_visible(), reason="a GPU is present; covered by the gpu tests")
def test_no_gpu_is_an_error_not_a_crash():
    L = F.load_library()
    h = C.c_void_p()
    rc = L.fa_create(0, C.byref(h))
    assert rc == -4
    assert b"No supported GPU" in L.fa_last_error()
    with pytest.raises(F.FaError):
        F.Context(0)


def test_null_arguments_are_rejected():
    L = F.load_library()
    assert L.fa_set_reference(None, None, 1, 2, 2, 3, None, 0, None) == -1
    assert L.fa_register_frame(None, None, 0, 0.0, None) == -1
    assert L.fa_get_error(None, None, None) == -1
    assert L.fa_batch_run(None, None, None, None) == -1
