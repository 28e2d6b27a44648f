"""Simplex meshes and nodal DG discretizations: the part of hedge the particle path reads.

The reference builds its mesh contract from a hedge discretization
(pyrticle/meshdata.py:37-133) and its monomial pusher from hedge's local
discretization (pyrticle/pusher.py:87-114: ``node_tuples()``, ``unit_nodes()``,
``generic_vandermonde``).  hedge is not available, so this module rebuilds those
pieces from their published definitions:

* unit simplex with vertices at -1 (hedge convention): triangle
  (-1,-1),(1,-1),(-1,1); tetrahedron (-1,-1,-1),(1,-1,-1),(-1,1,-1),(-1,-1,1);
* warp-and-blend interpolation nodes (T. Warburton, "An explicit construction
  of interpolation nodes on the simplex", J. Eng. Math. 56, 2006; the Nodes2D /
  Nodes3D construction of Hesthaven & Warburton, Nodal DG Methods, 2008, s.6.1,
  s.10.1);
* affine element maps x = A r + b and their inverses, |det A|, outward face
  normals, face planes, centroids and bounding radii, face neighbours (with
  periodic partners) and vertex adjacency.

Only numpy/scipy are used.  Everything is float64.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field
from itertools import combinations

import numpy as np

INVALID_ELEMENT = 0xFFFFFFFF
INVALID_AXIS = 0xFFFFFFFF

# Local face -> local vertex table.  Face k of a simplex is the facet opposite
# to nothing in particular; we only need a fixed, documented order.
FACE_VERTICES = {
    1: [(0,), (1,)],
    2: [(0, 1), (1, 2), (0, 2)],
    3: [(0, 1, 2), (0, 1, 3), (0, 2, 3), (1, 2, 3)],
}

UNIT_VERTICES = {
    1: np.array([[-1.0], [1.0]]),
    2: np.array([[-1.0, -1.0], [1.0, -1.0], [-1.0, 1.0]]),
    3: np.array([[-1.0, -1.0, -1.0], [1.0, -1.0, -1.0], [-1.0, 1.0, -1.0],
                 [-1.0, -1.0, 1.0]]),
}


# ---------------------------------------------------------------------------
# 1D building blocks
# ---------------------------------------------------------------------------
def legendre_gauss_lobatto(n: int) -> np.ndarray:
    """The n+1 Legendre-Gauss-Lobatto points on [-1, 1], ascending."""
    if n == 0:
        return np.array([0.0])
    if n == 1:
        return np.array([-1.0, 1.0])
    interior = np.polynomial.legendre.Legendre.basis(n).deriv().roots()
    return np.concatenate([[-1.0], np.sort(interior.real), [1.0]])


def _warp_factor(n: int, rout: np.ndarray) -> np.ndarray:
    """Warburton's 1D warp function, divided by the edge blend 1 - r^2."""
    lgl = legendre_gauss_lobatto(n)
    req = np.linspace(-1.0, 1.0, n + 1)
    # Lagrange polynomials through the equispaced points, evaluated at rout.
    lag = np.ones((n + 1, rout.size))
    for i in range(n + 1):
        for j in range(n + 1):
            if i != j:
                lag[i] *= (rout - req[j]) / (req[i] - req[j])
    warp = lag.T @ (lgl - req)
    zerof = np.abs(rout) < 1.0 - 1.0e-10
    sf = 1.0 - (zerof * rout) ** 2
    return warp / sf + warp * (zerof - 1.0)


# ---------------------------------------------------------------------------
# warp & blend nodes
# ---------------------------------------------------------------------------
_ALPHA_OPT_2D = [0.0000, 0.0000, 1.4152, 0.1001, 0.2751, 0.9800, 1.0999, 1.2832,
                 1.3648, 1.4773, 1.4959, 1.5743, 1.5770, 1.6223, 1.6258]
_ALPHA_OPT_3D = [0.0, 0.0, 0.0, 0.1002, 1.1332, 1.5608, 1.3413, 1.2577, 1.1603,
                 1.10153, 0.6080, 0.4523, 0.8856, 0.8717, 0.9655]


def _equilateral_to_unit_triangle(x, y):
    l1 = (math.sqrt(3.0) * y + 1.0) / 3.0
    l2 = (-3.0 * x - math.sqrt(3.0) * y + 2.0) / 6.0
    l3 = (3.0 * x - math.sqrt(3.0) * y + 2.0) / 6.0
    return -l2 + l3 - l1, -l2 - l3 + l1


def warp_blend_nodes_2d(n: int) -> np.ndarray:
    """Warp & blend nodes on the unit triangle, shape ((n+1)(n+2)/2, 2)."""
    if n == 0:
        return np.array([[-1.0 / 3.0, -1.0 / 3.0]])
    alpha = _ALPHA_OPT_2D[n - 1] if n < 16 else 5.0 / 3.0
    lam1, lam3 = [], []
    for i in range(n + 1):
        for j in range(n + 1 - i):
            lam1.append(i / n)
            lam3.append(j / n)
    lam1 = np.array(lam1)
    lam3 = np.array(lam3)
    lam2 = 1.0 - lam1 - lam3
    x = -lam2 + lam3
    y = (-lam2 - lam3 + 2.0 * lam1) / math.sqrt(3.0)
    w1 = 4 * lam2 * lam3 * _warp_factor(n, lam3 - lam2) * (1 + (alpha * lam1) ** 2)
    w2 = 4 * lam1 * lam3 * _warp_factor(n, lam1 - lam3) * (1 + (alpha * lam2) ** 2)
    w3 = 4 * lam1 * lam2 * _warp_factor(n, lam2 - lam1) * (1 + (alpha * lam3) ** 2)
    c2, s2 = math.cos(2 * math.pi / 3), math.sin(2 * math.pi / 3)
    c4, s4 = math.cos(4 * math.pi / 3), math.sin(4 * math.pi / 3)
    x = x + w1 + c2 * w2 + c4 * w3
    y = y + s2 * w2 + s4 * w3
    r, s = _equilateral_to_unit_triangle(x, y)
    return np.stack([r, s], axis=1)


def _warp_shift_face_3d(n, alpha, l1, l2, l3):
    """In-face warp of the 2D construction, evaluated with barycentrics l1..l3.

    The edge blend 4*l2*l3 equals 1 - x^2 on an edge, so dividing the 1D warp by
    1 - x^2 (done by _warp_factor) makes edge nodes land on the LGL points.
    """
    wf1 = 4 * l2 * l3 * _warp_factor(n, l3 - l2) * (1 + (alpha * l1) ** 2)
    wf2 = 4 * l3 * l1 * _warp_factor(n, l1 - l3) * (1 + (alpha * l2) ** 2)
    wf3 = 4 * l1 * l2 * _warp_factor(n, l2 - l1) * (1 + (alpha * l3) ** 2)
    c2, s2 = math.cos(2 * math.pi / 3), math.sin(2 * math.pi / 3)
    c4, s4 = math.cos(4 * math.pi / 3), math.sin(4 * math.pi / 3)
    dx = wf1 + c2 * wf2 + c4 * wf3
    dy = s2 * wf2 + s4 * wf3
    return dx, dy


def warp_blend_nodes_3d(n: int) -> np.ndarray:
    """Warp & blend nodes on the unit tetrahedron, shape ((n+1)(n+2)(n+3)/6, 3)."""
    if n == 0:
        return np.array([[-0.5, -0.5, -0.5]])
    alpha = _ALPHA_OPT_3D[n - 1] if n < 16 else 1.0
    tol = 1.0e-10
    r, s, t = [], [], []
    for k in range(n + 1):
        for j in range(n + 1 - k):
            for i in range(n + 1 - k - j):
                r.append(-1.0 + 2.0 * i / n)
                s.append(-1.0 + 2.0 * j / n)
                t.append(-1.0 + 2.0 * k / n)
    r, s, t = map(np.array, (r, s, t))
    lam = [(1 + t) / 2, (1 + s) / 2, -(1 + r + s + t) / 2, (1 + r) / 2]
    v1 = np.array([-1.0, -1.0 / math.sqrt(3.0), -1.0 / math.sqrt(6.0)])
    v2 = np.array([1.0, -1.0 / math.sqrt(3.0), -1.0 / math.sqrt(6.0)])
    v3 = np.array([0.0, 2.0 / math.sqrt(3.0), -1.0 / math.sqrt(6.0)])
    v4 = np.array([0.0, 0.0, 3.0 / math.sqrt(6.0)])
    t1 = np.array([v2 - v1, v2 - v1, v3 - v2, v3 - v1])
    t2 = np.array([v3 - 0.5 * (v1 + v2), v4 - 0.5 * (v1 + v2),
                   v4 - 0.5 * (v2 + v3), v4 - 0.5 * (v1 + v3)])
    t1 /= np.linalg.norm(t1, axis=1)[:, None]
    t2 /= np.linalg.norm(t2, axis=1)[:, None]
    xyz = (np.outer(lam[2], v1) + np.outer(lam[3], v2) + np.outer(lam[1], v3)
           + np.outer(lam[0], v4))
    shift = np.zeros_like(xyz)
    orders = [(0, 1, 2, 3), (1, 0, 2, 3), (2, 0, 3, 1), (3, 0, 2, 1)]
    for face, (a, b, c, d) in enumerate(orders):
        la, lb, lc, ld = lam[a], lam[b], lam[c], lam[d]
        w1, w2 = _warp_shift_face_3d(n, alpha, lb, lc, ld)
        blend = lb * lc * ld
        denom = (lb + 0.5 * la) * (lc + 0.5 * la) * (ld + 0.5 * la)
        ids = denom > tol
        blend[ids] = (1 + (alpha * la[ids]) ** 2) * blend[ids] / denom[ids]
        shift += np.outer(blend * w1, t1[face]) + np.outer(blend * w2, t2[face])
        on_face = (la < tol) & (((lb > tol).astype(int) + (lc > tol) + (ld > tol)) < 3)
        shift[on_face] = (np.outer(w1[on_face], t1[face])
                          + np.outer(w2[on_face], t2[face]))
    xyz = xyz + shift
    # equilateral tetrahedron -> unit tetrahedron
    rhs = xyz.T - 0.5 * (v2 + v3 + v4 - v1)[:, None]
    a_mat = np.column_stack([0.5 * (v2 - v1), 0.5 * (v3 - v1), 0.5 * (v4 - v1)])
    return np.linalg.solve(a_mat, rhs).T


def unit_nodes(dim: int, order: int) -> np.ndarray:
    if dim == 1:
        return legendre_gauss_lobatto(order)[:, None]
    if dim == 2:
        return warp_blend_nodes_2d(order)
    if dim == 3:
        return warp_blend_nodes_3d(order)
    raise ValueError("dimension must be 1, 2 or 3")


def node_tuples(dim: int, order: int) -> list:
    """All exponent tuples with sum <= order, in graded lexicographic order.

    These are the exponents of the monomial basis r^i s^j t^k the monomial pusher
    uses (pyrticle/pusher.py:104-105 takes them from hedge's node_tuples()).
    """
    out = []
    for total in range(order + 1):
        if dim == 1:
            out.append((total,))
        elif dim == 2:
            for j in range(total + 1):
                out.append((total - j, j))
        else:
            for j in range(total + 1):
                for k in range(total - j + 1):
                    out.append((total - j - k, j, k))
    return out


def monomial_vandermonde(points: np.ndarray, exponents) -> np.ndarray:
    """V[i, k] = prod_d points[i, d] ** exponents[k][d] (generic_vandermonde)."""
    exps = np.asarray(exponents)
    return np.prod(points[:, None, :] ** exps[None, :, :], axis=2)


def _monomial_integral_unit_simplex(exps) -> float:
    """Exact integral of prod r_d^e_d over the unit simplex with vertices at -1.

    Substituting r = 2*lambda - 1 maps the standard simplex {lambda >= 0,
    sum lambda <= 1} onto it with Jacobian 2^d; the Dirichlet integral
    int prod lambda^a = prod a! / (sum a + d)! finishes the job.
    """
    dim = len(exps)
    total = 0.0
    ranges = [range(e + 1) for e in exps]
    import itertools
    for ks in itertools.product(*ranges):
        coef = 1.0
        for e, k in zip(exps, ks):
            coef *= math.comb(e, k) * (2.0 ** k) * ((-1.0) ** (e - k))
        num = 1.0
        for k in ks:
            num *= math.factorial(k)
        total += coef * num / math.factorial(sum(ks) + dim)
    return total * 2.0 ** dim


@dataclass
class LocalDiscretization:
    """hedge-shaped local discretization on the unit simplex."""

    dimensions: int
    order: int
    _nodes: np.ndarray = field(init=False, repr=False)

    def __post_init__(self):
        self._nodes = unit_nodes(self.dimensions, self.order)

    def unit_nodes(self) -> np.ndarray:
        return self._nodes

    def node_tuples(self) -> list:
        return node_tuples(self.dimensions, self.order)

    def node_count(self) -> int:
        return self._nodes.shape[0]

    def vandermonde(self) -> np.ndarray:
        return monomial_vandermonde(self._nodes, self.node_tuples())

    def integral_weights(self) -> np.ndarray:
        """w_i = int l_i over the unit simplex (M 1 with M the mass matrix)."""
        vdm = self.vandermonde()
        g = np.array([_monomial_integral_unit_simplex(e) for e in self.node_tuples()])
        return np.linalg.solve(vdm.T, g)

    def mass_matrix(self) -> np.ndarray:
        exps = self.node_tuples()
        gram = np.array([[_monomial_integral_unit_simplex(tuple(a + b for a, b in zip(ea, eb)))
                          for eb in exps] for ea in exps])
        vinv = np.linalg.inv(self.vandermonde())
        return vinv.T @ gram @ vinv


# ---------------------------------------------------------------------------
# mesh
# ---------------------------------------------------------------------------
class Element:
    """hedge-shaped view of one simplex of a :class:`Mesh` (built on demand)."""

    __slots__ = ("id", "vertex_indices", "map_matrix", "map_vector",
                 "inverse_matrix", "inverse_vector", "face_normals")

    def __init__(self, mesh, eid):
        self.id = int(eid)
        self.vertex_indices = tuple(int(v) for v in mesh.element_vertices[eid])
        self.map_matrix = mesh.map_matrix[eid]
        self.map_vector = mesh.map_vector[eid]
        self.inverse_matrix = mesh.inverse_matrix[eid]
        self.inverse_vector = mesh.inverse_vector[eid]
        self.face_normals = list(mesh.face_normals[eid])

    def jacobian(self) -> float:
        return float(np.linalg.det(self.map_matrix))

    def face_vertices(self, vertex_indices):
        dim = len(vertex_indices) - 1
        return [tuple(vertex_indices[i] for i in fv) for fv in FACE_VERTICES[dim]]


class _ElementList:
    def __init__(self, mesh):
        self._mesh = mesh

    def __len__(self):
        return self._mesh.element_vertices.shape[0]

    def __getitem__(self, i):
        if isinstance(i, slice):
            return [Element(self._mesh, k) for k in range(*i.indices(len(self)))]
        if i < 0:
            i += len(self)
        if not 0 <= i < len(self):
            raise IndexError(i)
        return Element(self._mesh, i)

    def __iter__(self):
        for k in range(len(self)):
            yield Element(self._mesh, k)


class Mesh:
    """A conforming simplex mesh with optional periodicity, stored as arrays.

    ``periodicity[axis]`` is None or a pair of boundary tags (as in hedge); the
    periodic extent is the bounding box along that axis.  Elements are
    reoriented to positive Jacobian.  Per element: the affine map x = A r + b
    taking unit vertex k to element vertex k, its inverse, and outward unit face
    normals in FACE_VERTICES order.
    """

    def __init__(self, points, elements, periodicity=None):
        points = np.ascontiguousarray(points, dtype=np.float64)
        self.points = points
        self.dimensions = dim = points.shape[1]
        ev = np.array(elements, dtype=np.int64).reshape(-1, dim + 1)
        verts = points[ev]
        edges = verts[:, 1:, :] - verts[:, :1, :]
        det = np.linalg.det(np.transpose(edges, (0, 2, 1)))
        flip = det < 0
        ev[flip, 1], ev[flip, 2] = ev[flip, 2].copy(), ev[flip, 1].copy()
        verts = points[ev]
        self.element_vertices = ev
        a = np.transpose(verts[:, 1:, :] - verts[:, :1, :], (0, 2, 1)) / 2.0
        uv0 = UNIT_VERTICES[dim][0]
        b = verts[:, 0, :] - np.einsum("kij,j->ki", a, uv0)
        self.map_matrix = a
        self.map_vector = b
        self.inverse_matrix = np.linalg.inv(a)
        self.inverse_vector = -np.einsum("kij,kj->ki", self.inverse_matrix, b)
        self.jacobians = np.linalg.det(a)
        centroid = verts.mean(axis=1)
        normals = np.empty((ev.shape[0], dim + 1, dim))
        for f, fv in enumerate(FACE_VERTICES[dim]):
            fp = verts[:, list(fv), :]
            if dim == 2:
                t = fp[:, 1] - fp[:, 0]
                n = np.stack([t[:, 1], -t[:, 0]], axis=1)
            else:
                n = np.cross(fp[:, 1] - fp[:, 0], fp[:, 2] - fp[:, 0])
            n /= np.linalg.norm(n, axis=1)[:, None]
            sgn = np.einsum("ki,ki->k", n, fp.mean(axis=1) - centroid)
            n[sgn < 0] *= -1
            normals[:, f] = n
        self.face_normals = normals
        self.elements = _ElementList(self)
        self.periodicity = list(periodicity) if periodicity is not None else [None] * dim
        self._build_connectivity()

    def bounding_box(self):
        return self.points.min(axis=0), self.points.max(axis=0)

    def _build_connectivity(self):
        dim = self.dimensions
        K = self.element_vertices.shape[0]
        F = dim + 1
        lo, hi = self.bounding_box()
        fv = np.stack([self.element_vertices[:, list(f)] for f in FACE_VERTICES[dim]], axis=1)
        keys = np.sort(fv.reshape(K * F, dim), axis=1)
        order = np.lexsort(keys.T[::-1])
        sk = keys[order]
        same = np.all(sk[1:] == sk[:-1], axis=1)
        nb = np.full(K * F, INVALID_ELEMENT, dtype=np.int64)
        ax = np.full(K * F, INVALID_AXIS, dtype=np.int64)
        i0 = order[:-1][same]
        i1 = order[1:][same]
        nb[i0] = i1 // F
        nb[i1] = i0 // F
        matched = np.zeros(K * F, dtype=bool)
        matched[i0] = True
        matched[i1] = True
        boundary = np.nonzero(~matched)[0]
        self.periodic_opposite_faces = {}
        self.periodic_opposite_vertices = {}
        still_boundary = set(boundary.tolist())
        for axis, per in enumerate(self.periodicity):
            if per is None:
                continue
            period = hi[axis] - lo[axis]
            tol = 1e-9 * max(1.0, period)

            def key_of(face_flat, shift):
                pts = self.points[keys[face_flat]].copy()
                pts[:, axis] += shift
                return tuple(sorted(tuple(np.round(q / tol).astype(np.int64)) for q in pts))

            low = {}
            for fl in boundary:
                pts = self.points[keys[fl]]
                if np.all(np.abs(pts[:, axis] - lo[axis]) < tol):
                    low[key_of(fl, 0.0)] = fl
            for fl in boundary:
                pts = self.points[keys[fl]]
                if np.all(np.abs(pts[:, axis] - hi[axis]) < tol):
                    k = key_of(fl, -period)
                    if k in low:
                        lf = low[k]
                        nb[fl] = lf // F
                        nb[lf] = fl // F
                        ax[fl] = axis
                        ax[lf] = axis
                        a_key, b_key = tuple(keys[fl]), tuple(keys[lf])
                        self.periodic_opposite_faces[a_key] = (b_key, axis)
                        self.periodic_opposite_faces[b_key] = (a_key, axis)
                        still_boundary.discard(int(fl))
                        still_boundary.discard(int(lf))
            low_v = {}
            on_lo = np.nonzero(np.abs(self.points[:, axis] - lo[axis]) < tol)[0]
            on_hi = np.nonzero(np.abs(self.points[:, axis] - hi[axis]) < tol)[0]
            for vi in on_lo:
                low_v[tuple(np.round(self.points[vi] / tol).astype(np.int64))] = int(vi)
            for vi in on_hi:
                q = self.points[vi].copy()
                q[axis] -= period
                k = tuple(np.round(q / tol).astype(np.int64))
                if k in low_v:
                    lv = low_v[k]
                    self.periodic_opposite_vertices.setdefault(int(vi), []).append((lv, axis))
                    self.periodic_opposite_vertices.setdefault(lv, []).append((int(vi), axis))
        self.neighbors = nb.reshape(K, F).astype(np.uint64)
        self.neighbor_axis = ax.reshape(K, F).astype(np.uint64)
        self.boundary_faces = {frozenset(keys[fl].tolist()) for fl in still_boundary}


@dataclass
class ElementGroup:
    local_discretization: LocalDiscretization
    members: list


class Discretization:
    """Nodal DG discretization of a mesh with one uniform element group.

    Element ``k`` owns nodes ``[k*Np, (k+1)*Np)`` (the reference's interpolator
    indexes fields by ``en*Np``, push_monomial.hpp:108-109).
    """

    def __init__(self, mesh: Mesh, order: int):
        self.mesh = mesh
        self.order = order
        self.dimensions = mesh.dimensions
        ldis = LocalDiscretization(mesh.dimensions, order)
        self.element_groups = [ElementGroup(ldis, mesh.elements)]
        un = ldis.unit_nodes()
        self.np = un.shape[0]
        nodes = np.einsum("kij,nj->kni", mesh.map_matrix, un) + mesh.map_vector[:, None, :]
        self.nodes = np.ascontiguousarray(nodes.reshape(-1, self.dimensions))

    def __len__(self):
        return self.nodes.shape[0]

    def find_el_range(self, el_id):
        return range(el_id * self.np, (el_id + 1) * self.np)

    def volume_zeros(self):
        return np.zeros(len(self))

    def interpolate_volume_function(self, f):
        return np.array([f(x) for x in self.nodes])

    def integral(self, vec):
        w = self.element_groups[0].local_discretization.integral_weights()
        jac = np.abs(self.mesh.jacobians)
        v = np.asarray(vec).reshape(len(self.mesh.elements), self.np)
        return float(np.sum(jac[:, None] * w[None, :] * v))


# ---------------------------------------------------------------------------
# seeded mesh generators
# ---------------------------------------------------------------------------
def make_jittered_square_mesh(n: int = 33, lo: float = -1.0, hi: float = 1.0,
                              jitter: float = 0.25, seed: int = 0) -> Mesh:
    """Delaunay mesh of an n x n jittered grid on [lo, hi]^2.

    Interior points move by U(+-jitter*h) per axis, boundary points stay on the
    boundary (sliding along it), so the domain remains the exact square.
    """
    from scipy.spatial import Delaunay
    rng = np.random.default_rng(seed)
    h = (hi - lo) / (n - 1)
    g = np.linspace(lo, hi, n)
    xx, yy = np.meshgrid(g, g, indexing="ij")
    pts = np.stack([xx.ravel(), yy.ravel()], axis=1)
    jit = rng.uniform(-jitter * h, jitter * h, size=pts.shape)
    on_x = (np.abs(pts[:, 0] - lo) < 1e-12) | (np.abs(pts[:, 0] - hi) < 1e-12)
    on_y = (np.abs(pts[:, 1] - lo) < 1e-12) | (np.abs(pts[:, 1] - hi) < 1e-12)
    jit[on_x, 0] = 0.0
    jit[on_y, 1] = 0.0
    corner = on_x & on_y
    jit[corner] = 0.0
    pts = pts + jit
    tri = Delaunay(pts)
    return Mesh(pts, tri.simplices)


def _disk_points(radius: float, n_rings: int, rng, jitter: float) -> np.ndarray:
    pts = [np.zeros(2)]
    for k in range(1, n_rings + 1):
        r = radius * k / n_rings
        m = max(6, int(round(2 * math.pi * k)))
        phase = rng.uniform(0, 2 * math.pi / m) if k < n_rings else 0.0
        ang = phase + 2 * math.pi * np.arange(m) / m
        ring = np.stack([r * np.cos(ang), r * np.sin(ang)], axis=1)
        if k < n_rings and jitter > 0:
            ring += rng.uniform(-jitter, jitter, size=ring.shape) * radius / n_rings
        pts.append(ring)
    return np.vstack([pts[0][None, :]] + pts[1:])


def make_cylinder_mesh(radius: float = 1.0, z0: float = 0.0, z1: float = 4.0,
                       n_rings: int = 10, n_layers: int = 20, periodic: bool = False,
                       jitter: float = 0.15, seed: int = 0) -> Mesh:
    """Tetrahedral mesh of a cylinder along z, built by extrusion.

    A seeded ring-and-jitter point set on the disk is Delaunay-triangulated,
    extruded into ``n_layers`` prism layers, and every prism is split into three
    tetrahedra with each quadrilateral side cut along the diagonal that joins its
    lower-numbered bottom vertex to its higher-numbered top vertex, which keeps
    the split conforming across prisms.  The domain is convex.
    """
    from scipy.spatial import Delaunay
    rng = np.random.default_rng(seed)
    disk = _disk_points(radius, n_rings, rng, jitter)
    tri = Delaunay(disk).simplices
    n2 = disk.shape[0]
    zs = np.linspace(z0, z1, n_layers + 1)
    pts = np.empty(((n_layers + 1) * n2, 3))
    for l, z in enumerate(zs):
        pts[l * n2:(l + 1) * n2, :2] = disk
        pts[l * n2:(l + 1) * n2, 2] = z
    tets = []
    for l in range(n_layers):
        for t in tri:
            a, b, c = sorted(int(v) for v in t)
            bot = [a + l * n2, b + l * n2, c + l * n2]
            top = [a + (l + 1) * n2, b + (l + 1) * n2, c + (l + 1) * n2]
            tets.append([bot[0], bot[1], bot[2], top[2]])
            tets.append([bot[0], bot[1], top[1], top[2]])
            tets.append([bot[0], top[0], top[1], top[2]])
    per = [None, None, ("minus_z", "plus_z")] if periodic else None
    return Mesh(pts, np.array(tets), periodicity=per)


def make_box_mesh(lo, hi, n: int, jitter: float = 0.15, seed: int = 0) -> Mesh:
    """Tetrahedral mesh of an axis-aligned box from a jittered n^3 grid."""
    from scipy.spatial import Delaunay
    rng = np.random.default_rng(seed)
    lo = np.asarray(lo, float)
    hi = np.asarray(hi, float)
    axes = [np.linspace(lo[d], hi[d], n) for d in range(3)]
    g = np.stack(np.meshgrid(*axes, indexing="ij"), axis=-1).reshape(-1, 3)
    h = (hi - lo) / (n - 1)
    jit = rng.uniform(-jitter, jitter, size=g.shape) * h
    for d in range(3):
        on = (np.abs(g[:, d] - lo[d]) < 1e-12) | (np.abs(g[:, d] - hi[d]) < 1e-12)
        jit[on, d] = 0.0
    g = g + jit
    tet = Delaunay(g).simplices
    # drop degenerate slivers Delaunay may emit on the flat hull faces
    keep = []
    for t in tet:
        v = g[t]
        vol = abs(np.linalg.det(np.column_stack([v[1] - v[0], v[2] - v[0], v[3] - v[0]]))) / 6
        if vol > 1e-12 * np.prod(h):
            keep.append(t)
    return Mesh(g, np.array(keep))


def min_edge_length(mesh: Mesh) -> float:
    verts = mesh.points[mesh.element_vertices]
    best = np.inf
    for i, j in combinations(range(verts.shape[1]), 2):
        best = min(best, float(np.min(np.linalg.norm(verts[:, i] - verts[:, j], axis=1))))
    return best
