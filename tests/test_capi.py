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
