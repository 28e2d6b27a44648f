"""Device-resident mesh handle and the particle-path calls on torch tensors.

Every method forwards to one C-ABI entry point of libpyrticle_mi355x.so on
torch's current stream.  Particle arrays are torch CUDA tensors in the
reference's layout (positions/momenta n x d float64, charges/masses n float64,
containing elements n uint32 stored as int32 bit patterns).
"""

from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from .meshdata import MeshArrays

INVALID_ELEMENT = _lib.INVALID_ELEMENT
FP64 = _lib.PYR_FP64
FP32C = _lib.PYR_FP32C


def _host(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


def element_tensor(elem_np, device="cuda") -> torch.Tensor:
    """uint32 element ids -> device int32 tensor with the same bits."""
    a = np.ascontiguousarray(elem_np, dtype=np.uint32).view(np.int32)
    return torch.from_numpy(a.copy()).to(device)


def element_numpy(elem_t: torch.Tensor) -> np.ndarray:
    return elem_t.detach().cpu().numpy().view(np.uint32)


class DeviceMesh:
    """Owns one pyr_mesh handle (mesh contract in HBM)."""

    def __init__(self, ma: MeshArrays):
        self.ma = ma
        self.dim = ma.dim
        self.K = ma.K
        self.Np = ma.Np
        L = _lib.lib()
        keep = {}

        def p(name, arr, dt):
            a = _host(arr, dt)
            keep[name] = a
            return a.ctypes.data

        desc = _lib.MeshDesc(
            dim=ma.dim, K=ma.K, Np=ma.Np, nv=ma.nv, nadj=int(len(ma.vadj_elements)),
            inv_map=p("inv_map", ma.inv_map, np.float64),
            jacobian=p("jacobian", ma.jacobian, np.float64),
            el_vertices=p("el_vertices", ma.el_vertices, np.uint32),
            face_normal=p("face_normal", ma.face_normal, np.float64),
            face_rhs=p("face_rhs", ma.face_rhs, np.float64),
            face_centroid=p("face_centroid", ma.face_centroid, np.float64),
            face_radius=p("face_radius", ma.face_radius, np.float64),
            face_neighbor=p("face_neighbor", ma.face_neighbor, np.uint32),
            face_per_axis=p("face_per_axis", ma.face_per_axis, np.uint32),
            vertices=p("vertices", ma.vertices, np.float64),
            vadj_starts=p("vadj_starts", ma.vadj_starts, np.uint32),
            vadj_elements=p("vadj_elements", ma.vadj_elements if len(ma.vadj_elements)
                            else np.zeros(1, np.uint32), np.uint32),
            periodicity=p("periodicity", ma.periodicity, np.float64),
            nodes=p("nodes", ma.nodes, np.float64),
            exponents=p("exponents", ma.exponents, np.int32),
            vandermonde=p("vandermonde", ma.vandermonde, np.float64),
        )
        h = C.c_void_p()
        _lib.check(L.pyr_mesh_create(C.byref(desc), C.byref(h)))
        self._h = h
        self._lib = L

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                self._lib.pyr_mesh_destroy(h)
            except Exception:
                pass
            self._h = None

    @property
    def handle(self):
        return self._h

    def _field_ptrs(self, fields):
        nn = self.K * self.Np
        return (C.c_void_p * 6)(*[None if f is None else _lib.arg(f, "f64", nn, None, "field")
                                  for f in fields])

    def _state_ptrs(self, n, x, p, q, m, elem):
        return (_lib.arg(x, "f64", n, self.dim, "x"), _lib.arg(p, "f64", n, self.dim, "p"),
                _lib.arg(q, "f64", n, None, "q"), _lib.arg(m, "f64", n, None, "m"),
                _lib.arg(elem, "u32", n, None, "elem"))

    # -- lookup -----------------------------------------------------------
    def find_containing_element(self, x: torch.Tensor) -> torch.Tensor:
        _lib.sync_stream()
        n = x.shape[0]
        out = torch.empty(n, dtype=torch.int32, device=x.device)
        _lib.check(self._lib.pyr_find_containing_element(
            self._h, n, _lib.arg(x, "f64", n, self.dim, "x"), _lib.ptr(out)))
        return out

    def is_in_element(self, elem: torch.Tensor, x: torch.Tensor) -> torch.Tensor:
        _lib.sync_stream()
        n = x.shape[0]
        out = torch.empty(n, dtype=torch.uint8, device=x.device)
        _lib.check(self._lib.pyr_is_in_element(self._h, n, _lib.arg(elem, "u32", n, None, "elem"),
                                               _lib.arg(x, "f64", n, self.dim, "x"),
                                               _lib.ptr(out)))
        return out

    # -- particles ----------------------------------------------------------
    def locate(self, x, p, elem, counters=None):
        """Relocalize in place; returns the dead flags (uint8)."""
        _lib.sync_stream()
        n = x.shape[0]
        dead = torch.empty(n, dtype=torch.uint8, device=x.device)
        _lib.check(self._lib.pyr_locate(self._h, n, _lib.arg(x, "f64", n, self.dim, "x"),
                                        _lib.arg(p, "f64", n, self.dim, "p"),
                                        _lib.arg(elem, "u32", n, None, "elem"),
                                        _lib.ptr(dead), _lib.arg(counters, "i64", 4, None,
                                                                 "counters")))
        return dead

    def forces(self, fields, x, elem, q, v, vis=False):
        _lib.sync_stream()
        n, vd = v.shape
        F = torch.empty((n, vd), dtype=torch.float64, device=x.device)
        vis_t = [torch.empty((n, 3), dtype=torch.float64, device=x.device)
                 for _ in range(4)] if vis else [None] * 4
        fp = self._field_ptrs(fields)
        _lib.check(self._lib.pyr_forces(self._h, n, vd, fp, _lib.arg(x, "f64", n, self.dim, "x"),
                                        _lib.arg(elem, "u32", n, None, "elem"),
                                        _lib.arg(q, "f64", n, None, "q"), _lib.ptr(v), _lib.ptr(F),
                                        *[_lib.ptr(t) for t in vis_t]))
        if vis:
            return F, dict(zip(("pt_e", "pt_b", "el_force", "mag_force"), vis_t))
        return F

    def average_forces(self, radius, alpha, normalizer, fields, x, elem, q, v, integral_weights,
                       vis=False):
        """Shape-averaged Lorentz forces (averaging pusher).  Returns
        (F, integral, n_zero[, vis dict]); integral is the per-particle shape
        integral N sum_e |J_e| sum_i S_i w_i (0 for particles with no mass)."""
        _lib.sync_stream()
        n, vd = v.shape
        dev = x.device
        F = torch.empty((n, vd), dtype=torch.float64, device=dev)
        integral = torch.empty(n, dtype=torch.float64, device=dev)
        names = ("pt_e", "pt_b", "pt_e_stddev", "pt_b_stddev", "el_force", "mag_force")
        shapes = ((n, 3), (n, 3), (n,), (n,), (n, vd), (n, vd))
        vis_t = [torch.empty(sh, dtype=torch.float64, device=dev) for sh in shapes] \
            if vis else [None] * 6
        w = torch.as_tensor(np.asarray(integral_weights, dtype=np.float64), device=dev)
        assert w.shape == (self.Np,), "integral weights need one entry per element node"
        fp = self._field_ptrs(fields)
        nz = C.c_int64(0)
        _lib.check(self._lib.pyr_average_forces(
            self._h, radius, alpha, normalizer, n, vd, fp, _lib.arg(x, "f64", n, self.dim, "x"),
            _lib.arg(elem, "u32", n, None, "elem"), _lib.arg(q, "f64", n, None, "q"),
            _lib.ptr(v), _lib.ptr(w), _lib.ptr(F), _lib.ptr(integral),
            *[_lib.ptr(t) for t in vis_t], C.byref(nz)))
        if vis:
            return F, integral, int(nz.value), dict(zip(names, vis_t))
        return F, integral, int(nz.value)

    def set_static_fields(self, fields):
        _lib.sync_stream()
        _lib.check(self._lib.pyr_set_static_fields(self._h, self._field_ptrs(fields)))

    def lsrk4_steps(self, nsteps, dt, c, n, x, p, q, m, elem, ids=None, counters=None,
                    precision=FP64) -> int:
        _lib.sync_stream()
        nn = C.c_int64(n)
        _lib.check(self._lib.pyr_lsrk4_steps(self._h, precision, nsteps, dt, c, C.byref(nn),
                                             *self._state_ptrs(n, x, p, q, m, elem),
                                             _lib.arg(ids, "i64", n, None, "ids"),
                                             _lib.arg(counters, "i64", 4, None, "counters")))
        return int(nn.value)

    def lsrk4_steps_average(self, nsteps, dt, c, radius, alpha, normalizer, integral_weights,
                            n, x, p, q, m, elem, ids=None, counters=None) -> int:
        """LSRK4 steps with the averaging pusher in the static fields of
        set_static_fields; returns the new particle count."""
        _lib.sync_stream()
        w = torch.as_tensor(np.asarray(integral_weights, dtype=np.float64), device=x.device)
        nn = C.c_int64(n)
        _lib.check(self._lib.pyr_lsrk4_steps_average(
            self._h, nsteps, dt, c, radius, alpha, normalizer, _lib.ptr(w), C.byref(nn),
            *self._state_ptrs(n, x, p, q, m, elem), _lib.arg(ids, "i64", n, None, "ids"),
            _lib.arg(counters, "i64", 4, None, "counters")))
        return int(nn.value)

    def deposit(self, radius, alpha, normalizer, x, q, elem, v=None, want_rho=True,
                want_j=True, precision=FP64, rho=None, j=None):
        _lib.sync_stream()
        n, d = x.shape
        nn = self.K * self.Np
        if want_rho and rho is None:
            rho = torch.empty(nn, dtype=torch.float64, device=x.device)
        if want_j and v is not None and j is None:
            j = torch.empty((nn, d), dtype=torch.float64, device=x.device)
        _lib.check(self._lib.pyr_deposit(
            self._h, radius, alpha, normalizer, precision, n, d,
            _lib.arg(x, "f64", n, self.dim, "x"), _lib.arg(q, "f64", n, None, "q"),
            _lib.arg(v, "f64", n, d, "v") if v is not None else None,
            _lib.arg(elem, "u32", n, None, "elem"),
            _lib.arg(rho, "f64", nn, None, "rho") if want_rho else None,
            _lib.arg(j, "f64", nn, d, "j") if (want_j and v is not None) else None))
        return (rho if want_rho else None), (j if (want_j and v is not None) else None)

    def deposit_normalized(self, radius, alpha, normalizer, x, q, elem, integral_weights,
                           v=None, want_rho=True, want_j=True, precision=FP64, scale=None):
        """Normalized shape deposition (each particle's deposit integrates to its
        charge).  Returns (rho, j, n_zero): n_zero located particles had a zero
        shape integral and deposited nothing.  scale (device, n) receives q / I."""
        _lib.sync_stream()
        n, d = x.shape
        nn = self.K * self.Np
        rho = torch.empty(nn, dtype=torch.float64, device=x.device) if want_rho else None
        j = torch.empty((nn, d), dtype=torch.float64, device=x.device) \
            if (want_j and v is not None) else None
        w = torch.as_tensor(np.asarray(integral_weights, dtype=np.float64), device=x.device)
        assert w.shape == (self.Np,), "integral weights need one entry per element node"
        nz = C.c_int64(0)
        _lib.check(self._lib.pyr_deposit_normalized(
            self._h, radius, alpha, normalizer, precision, n, d,
            _lib.arg(x, "f64", n, self.dim, "x"), _lib.arg(q, "f64", n, None, "q"),
            _lib.arg(v, "f64", n, d, "v") if v is not None else None,
            _lib.arg(elem, "u32", n, None, "elem"), _lib.ptr(w),
            _lib.ptr(rho), _lib.ptr(j), _lib.arg(scale, "f64", n, None, "scale"), C.byref(nz)))
        return rho, j, int(nz.value)

    def face_walk_stats(self, radius, x, elem):
        """Per particle: (elements visited by the reference's face walk, and
        sum, sum of squares, min, max of their centroid distances)."""
        _lib.sync_stream()
        n = x.shape[0]
        n_el = torch.empty(n, dtype=torch.int32, device=x.device)
        out = [torch.empty(n, dtype=torch.float64, device=x.device) for _ in range(4)]
        _lib.check(self._lib.pyr_face_walk_stats(self._h, radius, n,
                                                 _lib.arg(x, "f64", n, self.dim, "x"),
                                                 _lib.arg(elem, "u32", n, None, "elem"),
                                                 _lib.ptr(n_el), *[_lib.ptr(t) for t in out]))
        return (n_el, *out)

    def deposit_backward(self, radius, alpha, normalizer, x, q, elem, rho_bar=None, j_bar=None,
                         v=None, precision=FP64):
        """(x_bar, v_bar) from the adjoint densities; v_bar is None without j_bar."""
        _lib.sync_stream()
        n, d = x.shape
        x_bar = torch.empty((n, d), dtype=torch.float64, device=x.device)
        v_bar = torch.empty((n, d), dtype=torch.float64, device=x.device) \
            if j_bar is not None else None
        _lib.check(self._lib.pyr_deposit_backward(
            self._h, radius, alpha, normalizer, precision, n, d,
            _lib.arg(x, "f64", n, self.dim, "x"), _lib.arg(q, "f64", n, None, "q"),
            _lib.arg(v, "f64", n, d, "v"), _lib.arg(elem, "u32", n, None, "elem"),
            _lib.arg(rho_bar, "f64", self.K * self.Np, None, "rho_bar"),
            _lib.arg(j_bar, "f64", self.K * self.Np, d, "j_bar"), _lib.ptr(x_bar),
            _lib.ptr(v_bar)))
        return x_bar, v_bar

    def lsrk4_forward(self, nsteps, dt, c, x, p, q, m, elem, death, ckpt_every=1,
                      ckpt_xp=None, ckpt_elem=None, counters=None, precision=FP64):
        """nsteps LSRK4 steps in place without compaction (dead particles keep
        their slots, death[i] = 5*step + stage); optional checkpoints."""
        _lib.sync_stream()
        n = x.shape[0]
        _lib.check(self._lib.pyr_lsrk4_forward(
            self._h, precision, nsteps, ckpt_every, dt, c, n,
            *self._state_ptrs(n, x, p, q, m, elem), _lib.arg(death, "u32", n, None, "death"),
            _lib.ptr(ckpt_xp), _lib.ptr(ckpt_elem),
            _lib.arg(counters, "i64", 4, None, "counters")))

    def lsrk4_adjoint(self, nsteps, ckpt_every, dt, c, q, m, death, ckpt_xp, ckpt_elem, x_bar,
                      p_bar, precision=FP64):
        """In place: final-state gradients -> initial-state gradients."""
        _lib.sync_stream()
        n = x_bar.shape[0]
        _lib.check(self._lib.pyr_lsrk4_adjoint(
            self._h, precision, nsteps, ckpt_every, dt, c, n, _lib.ptr(q), _lib.ptr(m),
            _lib.ptr(death), _lib.ptr(ckpt_xp), _lib.ptr(ckpt_elem), _lib.ptr(x_bar),
            _lib.ptr(p_bar)))

    def deposit_stats(self):
        arr = (C.c_ulonglong * 4)()
        _lib.check(self._lib.pyr_deposit_stats(self._h, C.byref(arr)))
        return tuple(int(v) for v in arr)

    def deposit_list_stats(self):
        """(list entries, listed chunk-group pairs, matrix-core instructions
        issued, work items) of the last FP32C deposition; zeros otherwise."""
        _lib.sync_stream()
        arr = (C.c_ulonglong * 4)()
        _lib.check(self._lib.pyr_deposit_list_stats(self._h, C.byref(arr)))
        return tuple(int(v) for v in arr)

    def candidate_pairs(self, radius) -> int:
        out = C.c_int64()
        _lib.check(self._lib.pyr_deposit_candidates(self._h, radius, C.byref(out)))
        return int(out.value)

    def sort_by_element(self, elem):
        _lib.sync_stream()
        n = elem.shape[0]
        perm = torch.empty(n, dtype=torch.int32, device=elem.device)
        seg = torch.empty(self.K + 1, dtype=torch.int32, device=elem.device)
        _lib.check(self._lib.pyr_sort_by_element(self._h, n, _lib.arg(elem, "u32", n, None,
                                                                      "elem"), _lib.ptr(perm),
                                                 _lib.ptr(seg)))
        return perm, seg

    def set_deposit_counting(self, on: bool):
        """FP64 deposition on this mesh also counts pair and visit statistics
        (read by deposit_stats)."""
        _lib.check(self._lib.pyr_deposit_set_counting(self._h, 1 if on else 0))

    def set_push_mfma(self, on: bool, count: bool = False):
        """FP32C push on this mesh: interpolate on the matrix cores for waves
        inside one element; with count, the particle-stages that take that path
        are counted."""
        _lib.check(self._lib.pyr_push_set_mfma(self._h, (2 if count else 1) if on else 0))

    def push_mfma_waves(self) -> int:
        """Particle-stages interpolated on the matrix cores since the last call."""
        out = C.c_int64(0)
        _lib.check(self._lib.pyr_push_mfma_waves(self._h, C.byref(out)))
        return int(out.value)


def velocities(p: torch.Tensor, m: torch.Tensor, c: float) -> torch.Tensor:
    _lib.sync_stream()
    n, vd = p.shape
    v = torch.empty_like(p)
    _lib.check(_lib.lib().pyr_velocities(n, vd, c, _lib.ptr(p), _lib.ptr(m), _lib.ptr(v)))
    return v


# particle-state diagnostics (device inputs; sums formed on the device)
def kinetic_energies(p: torch.Tensor, m: torch.Tensor, c: float) -> torch.Tensor:
    _lib.sync_stream()
    n, vd = p.shape
    out = torch.empty(n, dtype=torch.float64, device=p.device)
    _lib.check(_lib.lib().pyr_kinetic_energies(n, vd, c, _lib.ptr(p), _lib.ptr(m),
                                               _lib.ptr(out)))
    return out


def _host_out(k):
    return np.zeros(k, dtype=np.float64)


def total_charge(q: torch.Tensor) -> float:
    _lib.sync_stream()
    out = _host_out(1)
    _lib.check(_lib.lib().pyr_total_charge(q.shape[0], _lib.ptr(q), out.ctypes.data))
    return float(out[0])


def particle_momentum(p: torch.Tensor) -> np.ndarray:
    _lib.sync_stream()
    n, vd = p.shape
    out = _host_out(vd)
    _lib.check(_lib.lib().pyr_particle_momentum(n, vd, _lib.ptr(p), out.ctypes.data))
    return out


def particle_current(q: torch.Tensor, v: torch.Tensor, length: float) -> np.ndarray:
    _lib.sync_stream()
    n, vd = v.shape
    out = _host_out(vd)
    _lib.check(_lib.lib().pyr_particle_current(n, vd, _lib.ptr(q), _lib.ptr(v), length,
                                               out.ctypes.data))
    return out


def rms_beam_size(x: torch.Tensor, axis: int) -> float:
    _lib.sync_stream()
    n, xd = x.shape
    out = _host_out(1)
    _lib.check(_lib.lib().pyr_rms_beam_size(n, xd, _lib.ptr(x), axis, out.ctypes.data))
    return float(out[0])


def rms_beam_emittance(x: torch.Tensor, p: torch.Tensor, axis: int, beam_axis: int) -> float:
    _lib.sync_stream()
    n, xd = x.shape
    out = _host_out(1)
    _lib.check(_lib.lib().pyr_rms_beam_emittance(n, xd, p.shape[1], _lib.ptr(x), _lib.ptr(p),
                                                 axis, beam_axis, out.ctypes.data))
    return float(out[0])


def rms_energy_spread(p: torch.Tensor, m: torch.Tensor, c: float) -> float:
    _lib.sync_stream()
    n, vd = p.shape
    out = _host_out(1)
    _lib.check(_lib.lib().pyr_rms_energy_spread(n, vd, c, _lib.ptr(p), _lib.ptr(m),
                                                out.ctypes.data))
    return float(out[0])


def velocities_backward(p: torch.Tensor, m: torch.Tensor, c: float,
                        v_bar: torch.Tensor) -> torch.Tensor:
    """p_bar = (dv/dp)^T v_bar."""
    _lib.sync_stream()
    n, vd = p.shape
    p_bar = torch.empty_like(p)
    _lib.check(_lib.lib().pyr_velocities_backward(n, vd, c, _lib.ptr(p), _lib.ptr(m),
                                                  _lib.ptr(v_bar), _lib.ptr(p_bar)))
    return p_bar


def compact_plan(dead: torch.Tensor):
    """(n_new, moves_from, moves_to) for the reference's swap-remove kill order."""
    _lib.sync_stream()
    n = dead.shape[0]
    frm = torch.empty(max(n, 1), dtype=torch.int32, device=dead.device)
    to = torch.empty(max(n, 1), dtype=torch.int32, device=dead.device)
    n_new = C.c_int64()
    n_moves = C.c_int64()
    _lib.check(_lib.lib().pyr_compact_plan(n, _lib.ptr(dead), _lib.ptr(frm), _lib.ptr(to),
                                           C.byref(n_new), C.byref(n_moves)))
    k = int(n_moves.value)
    return int(n_new.value), frm[:k], to[:k]


def release_scratch():
    """Frees the library-level scratch of compact_plan (device memory)."""
    _lib.check(_lib.lib().pyr_release_scratch())


def apply_moves(buf: torch.Tensor, frm: torch.Tensor, to: torch.Tensor):
    """buf[to] = buf[frm] along the first axis, in place."""
    _lib.sync_stream()
    k = frm.shape[0]
    if k == 0:
        return
    row_bytes = buf.element_size() * (buf.numel() // max(1, buf.shape[0]))
    _lib.check(_lib.lib().pyr_apply_moves(_lib.ptr(buf), row_bytes, k, _lib.ptr(frm),
                                          _lib.ptr(to)))


PROFILE_REGIONS = {"push": 0, "gather": 1, "sort": 2, "compact": 3, "adjoint": 4,
                   "deposit_adjoint": 5, "integral": 6, "average": 7, "diagnostics": 8}


def profile_enable(on: bool = True):
    _lib.check(_lib.lib().pyr_profile_enable(1 if on else 0))


def profile_read(region: str):
    """(total milliseconds, launches) recorded for a region since enabling."""
    ms = C.c_double()
    cnt = C.c_int64()
    _lib.check(_lib.lib().pyr_profile_read(PROFILE_REGIONS[region], C.byref(ms), C.byref(cnt)))
    return float(ms.value), int(cnt.value)
