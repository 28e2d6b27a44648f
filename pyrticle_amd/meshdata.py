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

from .discretization import FACE_VERTICES, INVALID_AXIS, INVALID_ELEMENT, Discretization


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
    ev = mesh.element_vertices
    K = ev.shape[0]
    F = d + 1
    lo, hi = mesh.bounding_box()

    periodicity = np.zeros((d, 2))
    for axis, per in enumerate(mesh.periodicity):
        if per is not None:
            periodicity[axis] = (lo[axis], hi[axis])

    inv_map = np.concatenate([mesh.inverse_matrix.reshape(K, d * d), mesh.inverse_vector],
                             axis=1)
    pts = mesh.points
    fn = mesh.face_normals
    fvi = np.stack([ev[:, list(f)] for f in FACE_VERTICES[d]], axis=1)   # K x F x d
    fpts = pts[fvi]                                                       # K x F x d x d
    rhs_each = np.einsum("kfvi,kfi->kfv", fpts, fn)
    spread = rhs_each.max(axis=2) - rhs_each.min(axis=2)
    assert np.all(spread <= 1e-12 * np.maximum(1.0, np.abs(rhs_each).max(axis=2))), \
        "non-planar face"
    frhs = rhs_each.mean(axis=2)
    fcen = fpts.mean(axis=2)
    frad = np.linalg.norm(fpts - fcen[:, :, None, :], axis=3).max(axis=2)

    # vertex -> element adjacency (meshdata.py:108-130), sorted by (element, axis)
    v_ids = ev.ravel()
    e_ids = np.repeat(np.arange(K), F)
    a_ids = np.full(v_ids.shape, INVALID_AXIS, dtype=np.int64)
    extra_v, extra_e, extra_a = [], [], []
    for vi, partners in mesh.periodic_opposite_vertices.items():
        els = np.nonzero((ev == vi).any(axis=1))[0]
        for other, axis in partners:
            extra_v.extend([other] * len(els))
            extra_e.extend(els.tolist())
            extra_a.extend([axis] * len(els))
    v_all = np.concatenate([v_ids, np.asarray(extra_v, dtype=np.int64)])
    e_all = np.concatenate([e_ids, np.asarray(extra_e, dtype=np.int64)])
    a_all = np.concatenate([a_ids, np.asarray(extra_a, dtype=np.int64)])
    triples = np.unique(np.stack([v_all, e_all, a_all], axis=1), axis=0)  # sorted v, e, a
    counts = np.bincount(triples[:, 0], minlength=len(pts))
    starts = np.zeros(len(pts) + 1, dtype=np.uint32)
    starts[1:] = np.cumsum(counts)

    ldis = discr.element_groups[0].local_discretization
    exps = np.asarray(ldis.node_tuples(), dtype=np.int32)
    return MeshArrays(
        dim=d, K=K, Np=discr.np, nv=len(pts),
        inv_map=np.ascontiguousarray(inv_map), jacobian=np.abs(mesh.jacobians),
        el_vertices=np.ascontiguousarray(ev, dtype=np.uint32),
        face_normal=np.ascontiguousarray(fn), face_rhs=np.ascontiguousarray(frhs),
        face_centroid=np.ascontiguousarray(fcen), face_radius=np.ascontiguousarray(frad),
        face_neighbor=np.ascontiguousarray(mesh.neighbors, dtype=np.uint32),
        face_per_axis=np.ascontiguousarray(mesh.neighbor_axis, dtype=np.uint32),
        vertices=np.ascontiguousarray(pts, dtype=np.float64),
        vadj_starts=starts,
        vadj_elements=np.ascontiguousarray(triples[:, 1], dtype=np.uint32),
        vadj_axes=np.ascontiguousarray(triples[:, 2], dtype=np.uint32),
        periodicity=periodicity,
        nodes=np.ascontiguousarray(discr.nodes, dtype=np.float64),
        exponents=exps, vandermonde=ldis.vandermonde(),
    )


INVALID = INVALID_ELEMENT
