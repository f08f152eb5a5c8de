"""CPU checks of the C-ABI library: it loads, and it exports every function
include/matmri_hip.h declares (no compute: there is no GPU here)."""
import ctypes
import os
import re

import pytest

from conftest import ROOT, load_pkg

HEADER = os.path.join(ROOT, "include", "matmri_hip.h")


def _declared():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(mm_[a-z0-9_]+)\s*\(", src)))


def test_header_lists_entry_points():
    names = _declared()
    assert "mm_nufft_plan_create" in names and "mm_cgne_solve" in names
    assert len(names) >= 18


def test_library_exports_every_declared_symbol():
    mm = load_pkg()
    if not os.path.exists(mm.LIB_PATH):
        pytest.fail("libmatmri_hip.so is not built (run __graft_entry__.build())")
    L = ctypes.CDLL(mm.LIB_PATH)
    missing = [n for n in _declared() if not hasattr(L, n)]
    assert not missing, missing
    assert set(mm.EXPORTS) == set(_declared())


def test_no_device_is_reported_cleanly():
    mm = load_pkg()
    L = mm.lib()
    n = L.mm_device_count()
    assert n >= 0
    assert b"gfx950" in L.mm_version()
    if n == 0:
        h = ctypes.c_void_p()
        rc = L.mm_ctx_create(0, None, ctypes.byref(h))
        assert rc == 3  # MM_ERR_DEVICE
        assert b"device" in L.mm_last_error()
