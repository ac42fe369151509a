"""The C ABI library loads, exports every entry point its header declares and
rejects bad arguments before touching a GPU (CPU only: no compute calls)."""
import ctypes as C
import os
import re

import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "pic_mi355x.h")


def declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(pic_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol(pic):
    L = pic._lib.load()
    names = declared()
    assert len(names) > 30
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # the ctypes table covers the whole header
    table = set(pic._lib.SIGNATURES) | {"pic_last_error"}
    assert set(names) == table, set(names) ^ table


def test_library_is_the_in_tree_build(pic):
    L = pic._lib.load()
    assert os.path.dirname(pic._lib.LIB_PATH) == os.path.join(ROOT, "particleincell.jl_amd")
    assert L.pic_version() == 100


def _create(pic, nx, ny, xmin=0.0, xmax=1.0, ymin=0.0, ymax=1.0, cap=0):
    h = C.c_void_p()
    rc = pic._lib.load().pic_ctx_create(C.byref(h), 0, nx, ny, xmin, xmax, ymin, ymax, cap)
    return rc, h


@pytest.mark.parametrize("nx,ny", [(3, 16), (16, 2), (0, 16), (16, -8)])
def test_ctx_create_rejects_bad_mesh(pic, nx, ny):
    rc, h = _create(pic, nx, ny)
    assert rc == -1 and not h.value
    assert "at least 4" in pic._lib.last_error()


def test_ctx_create_rejects_empty_domain(pic):
    rc, _ = _create(pic, 16, 16, xmin=1.0, xmax=1.0)
    assert rc == -1 and "domain" in pic._lib.last_error()
    rc, _ = _create(pic, 16, 16, cap=-5)
    assert rc == -1 and "capacity" in pic._lib.last_error()


def test_null_arguments_are_errors(pic):
    L = pic._lib.load()
    assert L.pic_ctx_create(None, 0, 16, 16, 0.0, 1.0, 0.0, 1.0, 0) == -1
    assert L.pic_synchronize(None) == -1
    assert L.pic_compute_rho(None, 0) == -1
    assert L.pic_run(None, 0, 0.1, 1, 0, None) == -1
    assert L.pic_f90_push_x(None, None, None, None, None) == -1
    assert "null" in pic._lib.last_error()
    assert L.pic_ctx_destroy(None) == 0


def test_error_helper_raises(pic):
    with pytest.raises(pic._lib.PicError, match="at least 4"):
        pic._lib.call("pic_ctx_create", C.byref(C.c_void_p()), 0, 2, 16, 0.0, 1.0, 0.0, 1.0, 0)
