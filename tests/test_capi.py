"""The C-ABI library loads and exports every entry point that
include/ngcasa_gridding.h declares (no compute calls: runs without a GPU)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    with open(os.path.join(ROOT, "include", "ngcasa_gridding.h")) as f:
        text = f.read()
    return sorted(set(re.findall(r"^\s*(?:const\s+char\s*\*|int)\s*(ngc_\w+)\s*\(", text, re.M)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    for s in ("ngc_standard_grid", "ngc_imaging_weight_degrid", "ngc_briggs_factors",
              "ngc_aperture_grid", "ngc_grid_to_image", "ngc_normalize_mosaic",
              "ngc_context_create", "ngc_context_destroy", "ngc_last_error"):
        assert s in syms


def test_library_exports_every_declared_symbol():
    from cngi_prototype_amd import _native
    if not os.path.exists(_native.LIB_PATH):
        pytest.fail("libngcasa_hip.so is not built; run __graft_entry__.build()")
    lib = ctypes.CDLL(_native.LIB_PATH)
    for s in declared_symbols():
        assert hasattr(lib, s), s
    assert set(declared_symbols()) == set(_native.EXPORTS)


def test_version_and_error_string_without_gpu():
    from cngi_prototype_amd import _native
    lib = _native.load_library()
    assert b"gfx950" in lib.ngc_version()
    assert lib.ngc_last_error() is not None


def test_binding_covers_the_header():
    """The ctypes binding declares a prototype for every entry point of the
    header, and its dtype/mode/precision codes are the header's."""
    from cngi_prototype_amd import _native
    assert sorted(_native.EXPORTS) == declared_symbols()
    with open(os.path.join(ROOT, "include", "ngcasa_gridding.h")) as f:
        defines = dict((k, int(v)) for k, v in
                       re.findall(r"^#define (NGC_\w+) (-?\d+)", f.read(), re.M))
    for name in ("C64", "C128", "F32", "F64", "I32", "I64", "BOOL"):
        assert getattr(_native, name) == defines["NGC_" + name], name
    for name in ("IMAGE", "PSF", "DENSITY", "WEIGHT"):
        assert getattr(_native, "GRID_" + name) == defines["NGC_GRID_" + name], name


def _zero_args(fn):
    out = []
    for t in fn.argtypes:
        if t in (ctypes.c_double, ctypes.c_float):
            out.append(0.0)
        elif t in (ctypes.c_int, ctypes.c_int32, ctypes.c_int64, ctypes.c_uint64):
            out.append(0)
        else:
            out.append(None)
    return out


def test_entry_points_reject_a_null_handle_without_a_gpu():
    """Every entry point that takes a context or plan returns NGC_ERR_ARG (-1)
    for a NULL one, with a message, before any device call; the destroy
    functions accept NULL, as free() does."""
    from cngi_prototype_amd import _native
    lib = _native.load_library()
    lib.ngc_last_error.restype = ctypes.c_char_p
    null_ok = {"ngc_context_destroy", "ngc_grid_plan_destroy"}
    skip = {"ngc_context_create", "ngc_last_error", "ngc_version"}
    checked = 0
    for name in _native.EXPORTS:
        if name in skip:
            continue
        fn = getattr(lib, name)
        rc = fn(*_zero_args(fn))
        if name in null_ok:
            assert rc == 0, name
        else:
            assert rc == -1, name
            assert lib.ngc_last_error(), name
        checked += 1
    assert checked == len(_native.EXPORTS) - len(skip)


def test_context_create_rejects_a_null_out_pointer():
    from cngi_prototype_amd import _native
    lib = _native.load_library()
    lib.ngc_last_error.restype = ctypes.c_char_p
    assert lib.ngc_context_create(0, None) == -1
    assert b"NULL" in lib.ngc_last_error()
