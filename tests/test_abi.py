"""The C-ABI library loads on a CPU-only machine and exports every symbol the
header declares, with the ctypes signatures the Python layer binds."""

import ctypes as C
import os
import re

from pyrticle_amd import _lib


def test_library_exports_every_header_symbol():
    L = _lib.load_library()
    syms = _lib.header_symbols()
    assert len(syms) >= 20
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing


def test_python_bindings_cover_the_header():
    assert set(_lib.header_symbols()) == set(_lib._SIGNATURES)


def test_header_argument_counts_match_bindings():
    hdr = open(os.path.join(os.path.dirname(_lib.__file__), "..", "include",
                            "pyrticle_mi355x.h")).read()
    hdr = re.sub(r"/\*.*?\*/", "", hdr, flags=re.S)
    for name, (_res, args) in _lib._SIGNATURES.items():
        m = re.search(r"\b%s\s*\(([^)]*)\)" % name, hdr)
        assert m, name
        params = [p for p in m.group(1).split(",") if p.strip() and p.strip() != "void"]
        assert len(params) == len(args), (name, params, args)


def test_error_reporting_without_gpu():
    L = _lib.load_library()
    # argument validation happens before any device work
    rc = L.pyr_velocities(10, 4, 1.0, None, None, None)
    assert rc == _lib.PYR_ERR_INVALID_ARG
    assert b"invalid argument" in L.pyr_last_error()


def test_mesh_desc_layout_matches_header():
    # five int32 then sixteen pointers
    assert C.sizeof(_lib.MeshDesc) == 5 * 4 + 4 + 16 * 8


def test_argument_types_checked_before_the_call():
    """Wrong element types or short arrays never reach a kernel: the checked
    addresses raise ValueError first (the dtype and length checks run before
    the device check, so CPU tensors exercise them)."""
    import pytest
    import torch
    with pytest.raises(ValueError, match="dtype"):
        _lib.arg(torch.zeros(8, dtype=torch.float32), "f64", 8)
    with pytest.raises(ValueError, match="dtype"):
        _lib.arg(torch.zeros(8, dtype=torch.int64), "u32", 8)
    with pytest.raises(ValueError, match="shape"):
        _lib.arg(torch.zeros((8, 2), dtype=torch.float64), "f64", 8, 3)
    with pytest.raises(ValueError, match="rows"):
        _lib.arg(torch.zeros((4, 3), dtype=torch.float64), "f64", 8, 3)
    with pytest.raises(ValueError, match="entries"):
        _lib.arg(torch.zeros(4, dtype=torch.float64), "f64", 8)
    with pytest.raises(ValueError, match="device"):
        _lib.arg(torch.zeros((8, 3), dtype=torch.float64), "f64", 8, 3)
    assert _lib.arg(None, "f64", 8) is None
