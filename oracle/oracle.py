"""ctypes front end of the CPU restatement in oracle.c.

TEST INFRASTRUCTURE ONLY: tests/, ``__graft_entry__.smoke()`` and the
benchmark's CPU-baseline leg import this module as the checker.  The shipped
package never imports it.
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle_pyr.so")
INVALID = 0xFFFFFFFF


def build(force: bool = False) -> str:
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-C", _HERE, "-B" if force else "liboracle_pyr.so"],
                       check=True, stdout=subprocess.DEVNULL)
    return _LIB_PATH


class _Mesh(C.Structure):
    _fields_ = [("dim", C.c_int), ("K", C.c_int), ("Np", C.c_int), ("nv", C.c_int)] + [
        (name, C.c_void_p) for name in (
            "inv_map", "jacobian", "el_vertices", "face_normal", "face_rhs",
            "face_centroid", "face_radius", "face_neighbor", "face_per_axis",
            "vertices", "vadj_starts", "vadj_elements", "periodicity", "nodes")]


class _Counters(C.Structure):
    _fields_ = [("same", C.c_uint64), ("neighbor", C.c_uint64),
                ("vertex", C.c_uint64), ("global_", C.c_uint64)]

    def as_tuple(self):
        return (self.same, self.neighbor, self.vertex, self.global_)


_lib = None


def lib():
    global _lib
    if _lib is None:
        build()
        _lib = C.CDLL(_LIB_PATH)
        _lib.orc_shape_normalizer.restype = C.c_double
        _lib.orc_shape_value.restype = C.c_double
        _lib.orc_find_containing_element.restype = C.c_uint32
        _lib.orc_find_new_containing_element.restype = C.c_uint32
        _lib.orc_last_error.restype = C.c_char_p
        for name in ("orc_total_charge", "orc_rms_beam_size", "orc_rms_beam_emittance",
                     "orc_rms_energy_spread"):
            getattr(_lib, name).restype = C.c_double
    return _lib


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


class OracleMesh:
    """Holds the contiguous arrays of a MeshArrays and the C view of them."""

    def __init__(self, ma):
        self.ma = ma
        self._keep = {}
        fields = {}
        for name, dt in (("inv_map", np.float64), ("jacobian", np.float64),
                         ("el_vertices", np.uint32), ("face_normal", np.float64),
                         ("face_rhs", np.float64), ("face_centroid", np.float64),
                         ("face_radius", np.float64), ("face_neighbor", np.uint32),
                         ("face_per_axis", np.uint32), ("vertices", np.float64),
                         ("vadj_starts", np.uint32), ("vadj_elements", np.uint32),
                         ("periodicity", np.float64), ("nodes", np.float64)):
            arr = np.ascontiguousarray(getattr(ma, name), dtype=dt)
            self._keep[name] = arr
            fields[name] = arr.ctypes.data
        self.c = _Mesh(ma.dim, ma.K, ma.Np, ma.nv, **fields)
        vt = np.ascontiguousarray(ma.vandermonde.T)      # V^T, as pusher.py:107-109
        self.lu = np.asfortranarray(vt.copy())
        self.ipiv = np.zeros(ma.Np, dtype=np.int32)
        info = lib().orc_getrf(ma.Np, _ptr(self.lu), _ptr(self.ipiv))
        if info != 0:
            raise RuntimeError("singular matrix in getrf")
        self.exps = np.ascontiguousarray(ma.exponents, dtype=np.int32)

    @property
    def ref(self):
        return C.byref(self.c)


def is_in_unit_simplex(d, r, tol=1e-10):
    r = np.ascontiguousarray(r, dtype=np.float64)
    return bool(lib().orc_is_in_unit_simplex(d, _ptr(r), C.c_double(tol)))


def inverse_map(om, en, x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    r = np.empty(om.ma.dim)
    lib().orc_inverse_map(om.ref, C.c_uint32(en), _ptr(x), _ptr(r))
    return r


def find_containing_element(om, x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    return int(lib().orc_find_containing_element(om.ref, _ptr(x)))


def find_new_containing_element(om, x, p, prev):
    x = np.ascontiguousarray(x, dtype=np.float64)
    p = np.ascontiguousarray(p, dtype=np.float64)
    cnt = _Counters()
    err = C.c_int(0)
    r = lib().orc_find_new_containing_element(om.ref, _ptr(x), _ptr(p), C.c_uint32(prev),
                                              C.byref(cnt), C.byref(err))
    if err.value:
        raise RuntimeError(lib().orc_last_error().decode())
    return int(r), cnt.as_tuple()


def velocities(p, m, c):
    p = np.ascontiguousarray(p, dtype=np.float64)
    m = np.ascontiguousarray(m, dtype=np.float64)
    n, vd = p.shape
    v = np.empty_like(p)
    lib().orc_velocities(C.c_int64(n), vd, C.c_double(c), _ptr(p), _ptr(m), _ptr(v))
    return v


def shape_normalizer(radius, dims, alpha):
    return lib().orc_shape_normalizer(C.c_double(radius), dims, C.c_double(alpha))


def shape_value(radius, dims, alpha, r):
    r = np.ascontiguousarray(r, dtype=np.float64)
    nrm = shape_normalizer(radius, dims, alpha)
    return lib().orc_shape_value(C.c_double(nrm), C.c_double(radius), C.c_double(alpha),
                                 dims, _ptr(r))


def update_containing_elements(om, x, p, q, m, elem, extras=(), ids=None):
    """In-place relocalization with kills; returns (new count, counters)."""
    n = C.c_int64(x.shape[0])
    xd = x.shape[1]
    vd = p.shape[1]
    extra_arrays = [np.ascontiguousarray(e) for e in extras]
    for e, src in zip(extra_arrays, extras):
        assert e is src, "extras must be contiguous float64 arrays"
    ptrs = (C.c_void_p * max(1, len(extras)))(*[e.ctypes.data for e in extras])
    widths = (C.c_int * max(1, len(extras)))(*[e.shape[1] for e in extras])
    cnt = _Counters()
    rc = lib().orc_update_containing_elements(
        om.ref, xd, vd, C.byref(n), _ptr(x), _ptr(p), _ptr(q), _ptr(m), _ptr(elem),
        len(extras), ptrs, widths, _ptr(ids), C.byref(cnt))
    if rc:
        raise RuntimeError(lib().orc_last_error().decode())
    return int(n.value), cnt.as_tuple()


def deposit(om, radius, alpha, x, q, v=None, elem=None, want_rho=True, want_j=True):
    ma = om.ma
    x = np.ascontiguousarray(x, dtype=np.float64)
    q = np.ascontiguousarray(q, dtype=np.float64)
    elem = np.ascontiguousarray(elem, dtype=np.uint32)
    n, xd = x.shape
    vd = v.shape[1] if v is not None else xd
    if v is not None:
        v = np.ascontiguousarray(v, dtype=np.float64)
    rho = np.empty(ma.nnodes) if want_rho else None
    j = np.empty((ma.nnodes, vd)) if (want_j and v is not None) else None
    visits = np.zeros(3, dtype=np.uint64)
    rc = lib().orc_deposit(om.ref, C.c_double(radius), C.c_double(alpha), C.c_int64(n), xd,
                           vd, _ptr(x), _ptr(q), _ptr(v), _ptr(elem), _ptr(rho), _ptr(j),
                           _ptr(visits))
    if rc:
        raise RuntimeError("deposit failed")
    return rho, j, visits


def deposit_normalized(om, radius, alpha, x, q, v, elem, integral_weights, want_rho=True,
                       want_j=True):
    """dep_normshape.hpp:53-245: returns (rho, j, scale, n_zero), scale = q / I per
    particle (0 where the shape integral I vanishes)."""
    ma = om.ma
    x = np.ascontiguousarray(x, dtype=np.float64)
    q = np.ascontiguousarray(q, dtype=np.float64)
    elem = np.ascontiguousarray(elem, dtype=np.uint32)
    w = np.ascontiguousarray(integral_weights, dtype=np.float64)
    assert w.shape == (ma.Np,)
    n, xd = x.shape
    vd = v.shape[1] if v is not None else xd
    if v is not None:
        v = np.ascontiguousarray(v, dtype=np.float64)
    rho = np.empty(ma.nnodes) if want_rho else None
    j = np.empty((ma.nnodes, vd)) if (want_j and v is not None) else None
    scale = np.empty(n)
    nz = C.c_int64(0)
    rc = lib().orc_deposit_normalized(om.ref, C.c_double(radius), C.c_double(alpha),
                                      C.c_int64(n), xd, vd, _ptr(x), _ptr(q), _ptr(v),
                                      _ptr(elem), _ptr(w), _ptr(rho), _ptr(j), _ptr(scale),
                                      C.byref(nz))
    if rc:
        raise RuntimeError("normalized deposit failed")
    return rho, j, scale, int(nz.value)


def _field_ptrs(fields):
    arrs = [None if f is None else np.ascontiguousarray(f, dtype=np.float64) for f in fields]
    ptrs = (C.c_void_p * 6)(*[None if a is None else a.ctypes.data for a in arrs])
    return arrs, ptrs


def forces(om, x, elem, q, v, fields, vis=False):
    x = np.ascontiguousarray(x, dtype=np.float64)
    v = np.ascontiguousarray(v, dtype=np.float64)
    elem = np.ascontiguousarray(elem, dtype=np.uint32)
    q = np.ascontiguousarray(q, dtype=np.float64)
    n, xd = x.shape
    vd = v.shape[1]
    keep, ptrs = _field_ptrs(fields)
    f = np.empty((n, vd))
    vis_arrays = [np.empty((n, 3)) for _ in range(4)] if vis else [None] * 4
    rc = lib().orc_forces(om.ref, om.ma.Np, _ptr(om.exps), _ptr(om.lu), _ptr(om.ipiv),
                          C.c_int64(n), xd, vd, _ptr(x), _ptr(elem), _ptr(q), _ptr(v), ptrs,
                          _ptr(f), *[_ptr(a) for a in vis_arrays])
    if rc:
        raise RuntimeError("forces failed")
    if vis:
        return f, dict(zip(("pt_e", "pt_b", "el_force", "mag_force"), vis_arrays))
    return f


def face_walk_stats(om, radius, x, elem):
    """dep_normshape.hpp:131-136, 175-177 with element_finder.hpp:203-294:
    per particle (n_el, sum, sum of squares, min, max of the centroid
    distances of the visited elements)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    elem = np.ascontiguousarray(elem, dtype=np.uint32)
    n, xd = x.shape
    n_el = np.empty(n, dtype=np.uint32)
    out = [np.empty(n) for _ in range(4)]
    rc = lib().orc_face_walk_stats(om.ref, C.c_double(radius), C.c_int64(n), xd, _ptr(x),
                                   _ptr(elem), _ptr(n_el), *[_ptr(a) for a in out])
    if rc:
        raise RuntimeError("face walk failed")
    return (n_el, *out)


def average_forces(om, radius, alpha, x, q, v, elem, integral_weights, fields,
                   normalized=False, vis=False):
    """push_average.hpp:38-443 (averaging pusher over the plain or the
    normalized shape deposit): returns (force, charge, n_zero[, vis dict]);
    charge is the reconstructed charge per particle (0 where the targets saw
    none)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    v = np.ascontiguousarray(v, dtype=np.float64)
    elem = np.ascontiguousarray(elem, dtype=np.uint32)
    q = np.ascontiguousarray(q, dtype=np.float64)
    w = np.ascontiguousarray(integral_weights, dtype=np.float64)
    n, xd = x.shape
    vd = v.shape[1]
    keep, ptrs = _field_ptrs(fields)
    f = np.empty((n, vd))
    charge = np.empty(n)
    names = ("pt_e", "pt_b", "pt_e_stddev", "pt_b_stddev", "el_force", "mag_force")
    shapes = ((n, 3), (n, 3), (n,), (n,), (n, vd), (n, vd))
    vis_arrays = [np.empty(sh) for sh in shapes] if vis else [None] * 6
    nz = C.c_int64(0)
    rc = lib().orc_average_forces(om.ref, C.c_double(radius), C.c_double(alpha),
                                  int(normalized), C.c_int64(n), xd, vd, _ptr(x), _ptr(q),
                                  _ptr(v), _ptr(elem), _ptr(w), ptrs, _ptr(f), _ptr(charge),
                                  *[_ptr(a) for a in vis_arrays], C.byref(nz))
    if rc:
        raise RuntimeError("average forces failed")
    if vis:
        return f, charge, int(nz.value), dict(zip(names, vis_arrays))
    return f, charge, int(nz.value)


def lagrange_weights(om, en, x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    c = np.empty(om.ma.Np)
    lib().orc_lagrange_weights(om.ref, om.ma.Np, _ptr(om.exps), _ptr(om.lu), _ptr(om.ipiv),
                               C.c_uint32(en), _ptr(x), _ptr(c))
    return c


def lsrk4_coefficients():
    a = np.empty(5)
    b = np.empty(5)
    lib().orc_lsrk4_coefficients(_ptr(a), _ptr(b))
    return a, b


def lsrk4_run(om, fields, c_light, x, p, q, m, elem, dt, nsteps, ids=None):
    """Runs nsteps of the static-field particle RHS in place; returns (n, counters)."""
    keep, ptrs = _field_ptrs(fields)
    n = C.c_int64(x.shape[0])
    cnt = _Counters()
    rc = lib().orc_lsrk4_run(om.ref, om.ma.Np, _ptr(om.exps), _ptr(om.lu), _ptr(om.ipiv),
                             ptrs, C.c_double(c_light), x.shape[1], p.shape[1], C.byref(n),
                             _ptr(x), _ptr(p), _ptr(q), _ptr(m), _ptr(elem), _ptr(ids),
                             C.c_double(dt), nsteps, C.byref(cnt))
    if rc:
        raise RuntimeError(lib().orc_last_error().decode() or "lsrk4 failed")
    return int(n.value), cnt.as_tuple()


# ---------------------------------------------------------------------------
# diagnostics (diagnostics.hpp:37-183)
# ---------------------------------------------------------------------------
def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def kinetic_energies(p, m, c):
    p, m = _f64(p), _f64(m)
    out = np.empty(p.shape[0])
    lib().orc_kinetic_energies(C.c_int64(p.shape[0]), p.shape[1], C.c_double(c), _ptr(p),
                               _ptr(m), _ptr(out))
    return out


def total_charge(q):
    q = _f64(q)
    return lib().orc_total_charge(C.c_int64(q.shape[0]), _ptr(q))


def particle_momentum(p):
    p = _f64(p)
    out = np.empty(p.shape[1])
    lib().orc_particle_momentum(C.c_int64(p.shape[0]), p.shape[1], _ptr(p), _ptr(out))
    return out


def particle_current(q, v, length):
    q, v = _f64(q), _f64(v)
    out = np.empty(v.shape[1])
    lib().orc_particle_current(C.c_int64(v.shape[0]), v.shape[1], _ptr(q), _ptr(v),
                               C.c_double(length), _ptr(out))
    return out


def rms_beam_size(x, axis):
    x = _f64(x)
    return lib().orc_rms_beam_size(C.c_int64(x.shape[0]), x.shape[1], _ptr(x), axis)


def rms_beam_emittance(x, p, axis, beam_axis):
    x, p = _f64(x), _f64(p)
    return lib().orc_rms_beam_emittance(C.c_int64(x.shape[0]), x.shape[1], p.shape[1], _ptr(x),
                                        _ptr(p), axis, beam_axis)


def rms_energy_spread(p, m, c):
    p, m = _f64(p), _f64(m)
    return lib().orc_rms_energy_spread(C.c_int64(p.shape[0]), p.shape[1], C.c_double(c),
                                       _ptr(p), _ptr(m))
