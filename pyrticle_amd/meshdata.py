"""The mesh contract the particle path reads, and the drop-in ``MeshData`` class.

Restates pyrticle/meshdata.py:37-133 (``MeshData.fill_from_hedge``) against the
discretization shim in :mod:`pyrticle_amd.discretization`, producing flat
float64/uint32 arrays in the layout the HIP library uploads to HBM (and the
oracle reads on the host):

=================  ======================  =====================================
array              shape                   meaning (reference)
=================  ======================  =====================================
inv_map            K x (d*d + d)           inverse affine map A, b (r = A x + b)
jacobian           K                       |det| of the forward map
el_vertices        K x (d+1)               vertex ids (meshdata.hpp:98)
face_normal        K x (d+1) x d           outward unit normal (:76)
face_rhs           K x (d+1)               plane n.x = rhs (:83)
face_centroid      K x (d+1) x d           face bounding circle centre (:86)
face_radius        K x (d+1)               face bounding circle radius (:87)
face_neighbor      K x (d+1)               neighbour element or INVALID (:77)
face_per_axis      K x (d+1)               periodicity axis or INVALID (:78)
vertices           nv x d                  mesh vertices
vadj_starts        nv + 1                  vertex -> element CRS (:120-122)
vadj_elements      nadj
vadj_axes          nadj
periodicity        d x 2                   (min, max), min == max: none (:104-107)
nodes              K*Np x d                DG node coordinates
=================  ======================  =====================================

Vertex adjacency lists are sorted by (element id, axis); the reference stores a
Python set's iteration order (meshdata.py:108-130), which is not reproducible.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .discretization import INVALID_AXIS, INVALID_ELEMENT, Discretization


@dataclass
class MeshArrays:
    dim: int
    K: int
    Np: int
    nv: int
    inv_map: np.ndarray
    jacobian: np.ndarray
    el_vertices: np.ndarray
    face_normal: np.ndarray
    face_rhs: np.ndarray
    face_centroid: np.ndarray
    face_radius: np.ndarray
    face_neighbor: np.ndarray
    face_per_axis: np.ndarray
    vertices: np.ndarray
    vadj_starts: np.ndarray
    vadj_elements: np.ndarray
    vadj_axes: np.ndarray
    periodicity: np.ndarray
    nodes: np.ndarray
    exponents: np.ndarray      # nb x d monomial exponents (node tuples)
    vandermonde: np.ndarray    # Np x nb, V[i, k] = phi_k(unit node i)

    @property
    def nnodes(self) -> int:
        return self.K * self.Np


def build_mesh_arrays(discr: Discretization) -> MeshArrays:
    """Restates ``MeshData.fill_from_hedge`` (pyrticle/meshdata.py:37-133)."""
    mesh = discr.mesh
    d = mesh.dimensions
    K = len(mesh.elements)
    lo, hi = mesh.bounding_box()

    periodicity = np.zeros((d, 2))
    for axis, per in enumerate(mesh.periodicity):
        if per is not None:
            periodicity[axis] = (lo[axis], hi[axis])

    inv_map = np.empty((K, d * d + d))
    jac = np.empty(K)
    elv = np.empty((K, d + 1), dtype=np.uint32)
    fn = np.empty((K, d + 1, d))
    frhs = np.empty((K, d + 1))
    fcen = np.empty((K, d + 1, d))
    frad = np.empty((K, d + 1))
    fnb = np.asarray(mesh.neighbors, dtype=np.uint32)
    fax = np.asarray(mesh.neighbor_axis, dtype=np.uint32)
    pts = mesh.points
    for el in mesh.elements:
        i = el.id
        inv_map[i, :d * d] = el.inverse_matrix.ravel()
        inv_map[i, d * d:] = el.inverse_vector
        jac[i] = abs(el.jacobian())
        elv[i] = el.vertex_indices
        for f, fvi in enumerate(el.face_vertices(el.vertex_indices)):
            n = el.face_normals[f]
            fv = pts[list(fvi)]
            rhs = fv @ n
            assert np.ptp(rhs) <= 1e-12 * max(1.0, np.abs(rhs).max()), "non-planar face"
            fn[i, f] = n
            frhs[i, f] = rhs.mean()
            fcen[i, f] = fv.mean(axis=0)
            frad[i, f] = max(np.linalg.norm(p - fcen[i, f]) for p in fv)

    adj = [set() for _ in range(len(pts))]
    for el in mesh.elements:
        for vi in el.vertex_indices:
            adj[vi].add((el.id, INVALID_AXIS))
            for other, axis in mesh.periodic_opposite_vertices.get(vi, []):
                adj[other].add((el.id, axis))
    starts = np.zeros(len(pts) + 1, dtype=np.uint32)
    elems, axes = [], []
    for vi in range(len(pts)):
        for eid, axis in sorted(adj[vi]):
            elems.append(eid)
            axes.append(axis)
        starts[vi + 1] = len(elems)

    ldis = discr.element_groups[0].local_discretization
    exps = np.asarray(ldis.node_tuples(), dtype=np.int32)
    return MeshArrays(
        dim=d, K=K, Np=discr.np, nv=len(pts),
        inv_map=np.ascontiguousarray(inv_map), jacobian=jac,
        el_vertices=elv, face_normal=fn, face_rhs=frhs, face_centroid=fcen,
        face_radius=frad, face_neighbor=np.ascontiguousarray(fnb),
        face_per_axis=np.ascontiguousarray(fax),
        vertices=np.ascontiguousarray(pts, dtype=np.float64),
        vadj_starts=starts, vadj_elements=np.asarray(elems, dtype=np.uint32),
        vadj_axes=np.asarray(axes, dtype=np.uint32), periodicity=periodicity,
        nodes=np.ascontiguousarray(discr.nodes, dtype=np.float64),
        exponents=exps, vandermonde=ldis.vandermonde(),
    )


INVALID = INVALID_ELEMENT
