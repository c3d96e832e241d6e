"""The C-ABI library loads, exports every function declared in include/*.h,
and rejects bad arguments with status codes before any GPU call (these run
without a GPU)."""
import ctypes as C
import glob
import os
import re

import pytest

import amdflow
from amdflow import OpticalFlowError, TV_L1

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for path in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(path).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(of_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return sorted(names)


def test_headers_declare_the_reference_entry_points():
    names = declared_functions()
    for n in ["of_tvl1_create", "of_tvl1_run_u8", "of_tvl1_flow", "of_tvl1_push_f32",
              "of_tvl1_destroy", "of_hs_create", "of_hs_compute", "of_hs_destroy",
              "of_lk_create", "of_lk_push_u8", "of_lk_compute", "of_lk_destroy"]:
        assert n in names


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(amdflow.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_status_strings():
    for s in range(7):
        assert amdflow.status_string(s)
    assert "sigma" in amdflow.status_string(2)
    assert "gfx950" in amdflow.build_info()


@pytest.mark.parametrize("kw,status", [
    (dict(width=1, height=10), 1),            # too small
    (dict(width=64, height=64, theta=0.0), 1),
    (dict(width=64, height=64, zfactor=1.0), 1),
    (dict(width=64, height=64, warps=0), 1),
    (dict(width=4, height=40, nscales=1), 2),  # 5-tap pre-smoothing wider than the image
    (dict(width=40, height=4, nscales=1), 2),
])
def test_tvl1_create_rejects_bad_arguments(kw, status):
    w, h = kw.pop("width"), kw.pop("height")
    with pytest.raises(OpticalFlowError) as e:
        TV_L1(w, h, **kw)
    assert e.value.status == status


def test_hs_and_lk_create_reject_bad_arguments():
    lib = amdflow._lib
    h = C.c_void_p()
    assert lib.of_hs_create(0, 64, 64, 64, C.byref(h)) == 1
    assert lib.of_hs_create(1, 32, 64, 64, C.byref(h)) == 1       # stride < width
    assert lib.of_lk_create(4, 2, 5, 64, 64, C.byref(h)) == 6     # no 4-tap spatial filter
    assert lib.of_lk_create(3, 4, 5, 64, 64, C.byref(h)) == 6     # no 4-tap temporal filter
    assert lib.of_lk_create(3, 2, 4, 64, 64, C.byref(h)) == 1     # even window
    assert lib.of_tvl1_create(None, 64, 64, None) == 1            # no output handle
