"""The hedge-shaped discretization shim: nodes, weights, maps, connectivity."""

import numpy as np
import pytest

from pyrticle_amd import discretization as D
from pyrticle_amd.meshdata import build_mesh_arrays


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_nodes_and_weights(dim, order):
    ld = D.LocalDiscretization(dim, order)
    un = ld.unit_nodes()
    npts = (order + 1) * (order + 2) // 2 if dim == 2 else (order + 1) * (order + 2) * (order + 3) // 6
    assert un.shape == (npts, dim)
    assert np.all(un >= -1 - 1e-12) and np.all(un.sum(1) <= -(dim - 2) + 1e-12)
    assert len(ld.node_tuples()) == npts
    vol = 2.0 if dim == 2 else 4.0 / 3.0
    w = ld.integral_weights()
    assert abs(w.sum() - vol) < 1e-13
    # weights integrate every monomial of degree <= order exactly
    for e in ld.node_tuples():
        f = np.prod(un ** np.array(e), axis=1)
        assert abs(w @ f - D._monomial_integral_unit_simplex(e)) < 1e-12
    assert np.linalg.cond(ld.vandermonde()) < 1e5


@pytest.mark.parametrize("dim,order", [(2, 3), (2, 4), (3, 3), (3, 4)])
def test_edge_nodes_are_lgl(dim, order):
    un = D.unit_nodes(dim, order)
    on_edge = np.all(np.abs(un[:, 1:] + 1) < 1e-12, axis=1)
    assert np.allclose(np.sort(un[on_edge, 0]), D.legendre_gauss_lobatto(order), atol=1e-13)


def test_lgl_points_known_values():
    assert np.allclose(D.legendre_gauss_lobatto(3), [-1, -1 / np.sqrt(5), 1 / np.sqrt(5), 1])
    assert np.allclose(D.legendre_gauss_lobatto(4), [-1, -np.sqrt(3 / 7), 0, np.sqrt(3 / 7), 1])


def test_square_mesh_connectivity():
    m = D.make_jittered_square_mesh()
    assert len(m.elements) == 2048
    assert np.all(m.jacobians > 0)
    nb = m.neighbors.astype(np.int64)
    K = len(m.elements)
    for e in range(0, K, 37):
        for f in range(3):
            o = nb[e, f]
            if o != D.INVALID_ELEMENT:
                assert e in nb[o]
    assert (nb == D.INVALID_ELEMENT).sum() == 4 * 32     # boundary edges of a 33x33 grid


def test_cylinder_mesh_periodic_partners():
    m = D.make_cylinder_mesh(n_rings=3, n_layers=4, periodic=True)
    assert np.all(m.jacobians > 0)
    ax = m.neighbor_axis.astype(np.int64)
    assert np.all((ax == 2) | (ax == D.INVALID_AXIS))
    n_per = (ax == 2).sum()
    assert n_per > 0 and n_per % 2 == 0
    ma = build_mesh_arrays(D.Discretization(m, 1))
    assert ma.periodicity[2, 0] == 0.0 and ma.periodicity[2, 1] == 4.0


def test_mesh_arrays_face_geometry():
    m = D.make_cylinder_mesh(n_rings=3, n_layers=3)
    ma = build_mesh_arrays(D.Discretization(m, 2))
    # every vertex of a face lies on its plane, the opposite vertex inside
    for e in range(0, ma.K, 11):
        v = ma.vertices[ma.el_vertices[e]]
        for f, fv in enumerate(D.FACE_VERTICES[3]):
            n = ma.face_normal[e, f]
            assert np.allclose(v[list(fv)] @ n, ma.face_rhs[e, f], atol=1e-13)
            other = [k for k in range(4) if k not in fv][0]
            assert v[other] @ n < ma.face_rhs[e, f]
    # vertex adjacency: every element listed under each of its vertices
    for e in range(0, ma.K, 13):
        for vi in ma.el_vertices[e]:
            lst = ma.vadj_elements[ma.vadj_starts[vi]:ma.vadj_starts[vi + 1]]
            assert e in lst


def test_stats_gatherer_matches_numpy():
    """_internal.StatsGatherer (tools.hpp:340-400 stats_gatherer): one-at-a-time
    adds and batch merges give numpy's moments, extremes and count."""
    import numpy as np
    import torch
    from pyrticle_amd._internal import StatsGatherer
    rng = np.random.default_rng(0)
    a = rng.normal(3.0, 2.0, size=1000)
    one = StatsGatherer()
    for v in a:
        one.add(v)
    batched = StatsGatherer()
    batched.add_array(a[:10])
    batched.add_array(torch.from_numpy(a[10:600]))
    batched.add_array(a[600:])
    batched.add_array(np.zeros(0))
    for st in (one, batched):
        assert st.count() == a.size
        assert abs(st.mean() - a.mean()) < 1e-13
        assert abs(st.variance() / a.var() - 1) < 1e-12
        assert abs(st.variance_sample() / a.var(ddof=1) - 1) < 1e-12
        assert abs(st.standard_deviation_sample() / a.std(ddof=1) - 1) < 1e-12
        assert st.minimum() == a.min() and st.maximum() == a.max()
    one.reset()
    assert one.count() == 0
