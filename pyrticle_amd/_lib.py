"""ctypes binding of libpyrticle_mi355x.so (declared in include/pyrticle_mi355x.h).

This is the only way the Python layer reaches the particle path: there is no
CPU fallback.  If the library is missing or no GPU is visible, the first call
raises.
"""

from __future__ import annotations

import ctypes as C
import os
import threading

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpyrticle_mi355x.so")

PYR_OK = 0
PYR_ERR_INVALID_ARG = -1
PYR_ERR_HIP = -2
PYR_ERR_NO_BEST_NORMAL = -3
PYR_FP64 = 0
PYR_FP32C = 1
INVALID_ELEMENT = 0xFFFFFFFF

_P = C.c_void_p
_I64 = C.c_int64
_D = C.c_double


class MeshDesc(C.Structure):
    _fields_ = [("dim", C.c_int32), ("K", C.c_int32), ("Np", C.c_int32), ("nv", C.c_int32),
                ("nadj", C.c_int32)] + [(n, _P) for n in (
                    "inv_map", "jacobian", "el_vertices", "face_normal", "face_rhs",
                    "face_centroid", "face_radius", "face_neighbor", "face_per_axis",
                    "vertices", "vadj_starts", "vadj_elements", "periodicity", "nodes",
                    "exponents", "vandermonde")]


# name -> (restype, argtypes)
_SIGNATURES = {
    "pyr_version": (C.c_int, []),
    "pyr_init": (C.c_int, [C.c_int]),
    "pyr_last_error": (C.c_char_p, []),
    "pyr_set_stream": (C.c_int, [_P]),
    "pyr_sync": (C.c_int, []),
    "pyr_profile_enable": (C.c_int, [C.c_int]),
    "pyr_profile_read": (C.c_int, [C.c_int, C.POINTER(_D), C.POINTER(_I64)]),
    "pyr_mesh_create": (C.c_int, [C.POINTER(MeshDesc), C.POINTER(_P)]),
    "pyr_mesh_destroy": (C.c_int, [_P]),
    "pyr_find_containing_element": (C.c_int, [_P, _I64, _P, _P]),
    "pyr_is_in_element": (C.c_int, [_P, _I64, _P, _P, _P]),
    "pyr_velocities": (C.c_int, [_I64, C.c_int, _D, _P, _P, _P]),
    "pyr_locate": (C.c_int, [_P, _I64, _P, _P, _P, _P, _P]),
    "pyr_compact_plan": (C.c_int, [_I64, _P, _P, _P, C.POINTER(_I64), C.POINTER(_I64)]),
    "pyr_apply_moves": (C.c_int, [_P, _I64, _I64, _P, _P]),
    "pyr_release_scratch": (C.c_int, []),
    "pyr_compact_plan_static": (C.c_int, [_I64, _P, _P, _P, C.POINTER(_I64),
                                          C.POINTER(_I64)]),
    "pyr_forces": (C.c_int, [_P, _I64, C.c_int, _P * 6, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
    "pyr_set_static_fields": (C.c_int, [_P, _P * 6]),
    "pyr_lsrk4_steps": (C.c_int, [_P, C.c_int, C.c_int, _D, _D, C.POINTER(_I64), _P, _P, _P,
                                  _P, _P, _P, _P]),
    "pyr_deposit": (C.c_int, [_P, _D, _D, _D, C.c_int, _I64, C.c_int, _P, _P, _P, _P, _P,
                              _P]),
    "pyr_deposit_normalized": (C.c_int, [_P, _D, _D, _D, C.c_int, _I64, C.c_int, _P, _P, _P,
                                         _P, _P, _P, _P, _P, C.POINTER(_I64)]),
    "pyr_average_forces": (C.c_int, [_P, _D, _D, _D, _I64, C.c_int, _P, _P, _P, _P, _P, _P,
                                     _P, _P, _P, _P, _P, _P, _P, _P, C.POINTER(_I64)]),
    "pyr_kinetic_energies": (C.c_int, [_I64, C.c_int, _D, _P, _P, _P]),
    "pyr_total_charge": (C.c_int, [_I64, _P, _P]),
    "pyr_particle_momentum": (C.c_int, [_I64, C.c_int, _P, _P]),
    "pyr_particle_current": (C.c_int, [_I64, C.c_int, _P, _P, _D, _P]),
    "pyr_rms_beam_size": (C.c_int, [_I64, C.c_int, _P, C.c_int, _P]),
    "pyr_rms_beam_emittance": (C.c_int, [_I64, C.c_int, C.c_int, _P, _P, C.c_int, C.c_int, _P]),
    "pyr_rms_energy_spread": (C.c_int, [_I64, C.c_int, _D, _P, _P, _P]),
    "pyr_lsrk4_steps_average": (C.c_int, [_P, C.c_int, _D, _D, _D, _D, _D, _P,
                                          C.POINTER(_I64), _P, _P, _P, _P, _P, _P, _P]),
    "pyr_face_walk_stats": (C.c_int, [_P, _D, _I64, _P, _P, _P, _P, _P, _P, _P]),
    "pyr_deposit_set_counting": (C.c_int, [_P, C.c_int]),
    "pyr_push_set_mfma": (C.c_int, [_P, C.c_int]),
    "pyr_push_mfma_waves": (C.c_int, [_P, C.POINTER(_I64)]),
    "pyr_deposit_stats": (C.c_int, [_P, C.POINTER(C.c_ulonglong * 4)]),
    "pyr_deposit_list_stats": (C.c_int, [_P, C.POINTER(C.c_ulonglong * 4)]),
    "pyr_deposit_candidates": (C.c_int, [_P, _D, C.POINTER(_I64)]),
    "pyr_sort_by_element": (C.c_int, [_P, _I64, _P, _P, _P]),
    "pyr_lsrk4_forward": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, _D, _D, _I64, _P, _P, _P,
                                    _P, _P, _P, _P, _P, _P]),
    "pyr_lsrk4_adjoint": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, _D, _D, _I64, _P, _P, _P,
                                    _P, _P, _P, _P]),
    "pyr_velocities_backward": (C.c_int, [_I64, C.c_int, _D, _P, _P, _P, _P]),
    "pyr_deposit_backward": (C.c_int, [_P, _D, _D, _D, C.c_int, _I64, C.c_int, _P, _P, _P, _P,
                                       _P, _P, _P, _P]),
}

_lock = threading.Lock()
_lib = None
_initialized = False


def header_symbols() -> list:
    """Function names declared in include/pyrticle_mi355x.h."""
    import re
    hdr = os.path.join(HERE, "..", "include", "pyrticle_mi355x.h")
    with open(hdr) as f:
        text = f.read()
    return sorted(set(re.findall(r"\b(pyr_[a-z0-9_]+)\s*\(", text)))


def load_library():
    """Load the shared object without touching the GPU."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise RuntimeError(
                    "libpyrticle_mi355x.so is not built; run `python -m pyrticle_amd.build` "
                    "(there is no CPU fallback for the particle path)")
            lib = C.CDLL(LIB_PATH)
            for name, (res, args) in _SIGNATURES.items():
                fn = getattr(lib, name)
                fn.restype = res
                fn.argtypes = args
            _lib = lib
    return _lib


def lib():
    """The library, initialized on the current torch device."""
    global _initialized
    L = load_library()
    if not _initialized:
        import torch
        if not torch.cuda.is_available():
            raise RuntimeError("pyrticle_amd needs a HIP GPU: torch.cuda.is_available() is "
                               "False (there is no CPU fallback)")
        check(L.pyr_init(torch.cuda.current_device()))
        _initialized = True
    return L


def check(rc: int):
    if rc != PYR_OK:
        msg = _lib.pyr_last_error().decode(errors="replace") if _lib else "unknown error"
        if rc == PYR_ERR_INVALID_ARG:
            raise ValueError(msg)
        raise RuntimeError(msg)


def ptr(t) -> int | None:
    """Device address of a torch tensor (or None)."""
    if t is None:
        return None
    if not t.is_cuda:
        raise ValueError("expected a device tensor")
    if not t.is_contiguous():
        raise ValueError("expected a contiguous tensor")
    return t.data_ptr()


# element types the C ABI reads, by kind; u32 buffers may be int32 tensors
# holding the same bits
_KINDS = {"f64": ("float64",), "u32": ("int32", "uint32"), "i64": ("int64",),
          "u8": ("uint8", "bool")}
_DTYPE_NAMES = {}   # torch dtype -> its name


def arg(t, kind: str, rows: int | None = None, cols: int | None = None,
        name: str = "array") -> int | None:
    """Checked device address of a particle or mesh array (None stays None).

    kind: element type the C side reads ("f64", "u32", "i64", "u8"); rows: the
    entries the call touches along the first axis (the tensor may hold more);
    cols: the required second dimension (None: any shape whose total size
    covers rows).  A mismatch raises ValueError before any device code runs,
    so the kernel never reads a buffer with the wrong element size or length."""
    if t is None:
        return None
    want = _KINDS[kind]
    got = _DTYPE_NAMES.get(t.dtype)
    if got is None:
        got = _DTYPE_NAMES[t.dtype] = str(t.dtype).replace("torch.", "")
    if got not in want:
        raise ValueError("%s: expected dtype %s, got %s" % (name, " or ".join(want), got))
    if cols is not None:
        if t.dim() != 2 or t.shape[1] != cols:
            raise ValueError("%s: expected shape (n, %d), got %s" % (name, cols,
                                                                   tuple(t.shape)))
        if rows is not None and t.shape[0] < rows:
            raise ValueError("%s: expected at least %d rows, got %d" % (name, rows,
                                                                        t.shape[0]))
    elif rows is not None and t.numel() < rows:
        raise ValueError("%s: expected at least %d entries, got %d" % (name, rows, t.numel()))
    return ptr(t)


_stream_set = None


def sync_stream():
    """Run library work on torch's current stream (set once per stream change)."""
    global _stream_set
    import torch
    h = torch.cuda.current_stream().cuda_stream
    L = lib()
    if h != _stream_set:
        check(L.pyr_set_stream(C.c_void_p(h)))
        _stream_set = h
