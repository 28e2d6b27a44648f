"""Deposition paths at element orders beside P3 (P1, P2 and P4 tetrahedra on a
small cylinder): the FP32C and FP64 list gathers (whose grouped node tables pad
P2 and P4 elements to a multiple of four nodes: 12 and 36 slots), the
normalized deposition's shape integral and the averaging pusher's gather (whose
per-target node skip runs one lane per node, 35 for P4), each against the
oracle (dep_shape.hpp:67-150, dep_normshape.hpp:124-198, push_average.hpp:
338-443)."""

import numpy as np
import pytest
import torch

from oracle import oracle as O
from tests.conftest import rel_l2
from tests.test_gpu_kernels import (FP32C_TOL, FP64_TOL, _average_case, _dev,
                                    _normalized_case, _poly_fields)

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", params=[1, 2, 4])
def cylinder_order(request):
    from pyrticle_amd import workloads as W
    # 18,480 tets: elements of about twice R, where the FP32C form's expanded
    # distance keeps the rho error below 1e-6 (the error grows as (h / R)^2)
    discr, ma = W.cylinder_mesh_3d(n_rings=7, n_layers=20, order=request.param)
    return W.cylinder_3d(n_particles=6000, mesh=(discr, ma))


@pytest.fixture(scope="module")
def order_case(cylinder_order, cuda):
    from pyrticle_amd import device as dv
    wl = cylinder_order
    dm = dv.DeviceMesh(wl.ma)
    om = O.OracleMesh(wl.ma)
    x, q = wl.x, wl.q
    el = dv.element_numpy(dm.find_containing_element(_dev(x, cuda)))
    v = O.velocities(wl.p, wl.m, wl.c)
    return wl, dm, om, x, q, v, el


@pytest.mark.parametrize("precision", ["fp64", "fp32c"])
def test_deposit_orders_match_oracle(order_case, cuda, precision):
    from pyrticle_amd import device as dv
    wl, dm, om, x, q, v, el = order_case
    assert (el != dv.INVALID_ELEMENT).all()
    rho_o, j_o, _ = O.deposit(om, wl.radius, wl.alpha, x, q, v, el)
    nrm = O.shape_normalizer(wl.radius, 3, wl.alpha)
    prec = dv.FP64 if precision == "fp64" else dv.FP32C
    tol = FP64_TOL if precision == "fp64" else FP32C_TOL
    rho, j = dm.deposit(wl.radius, wl.alpha, nrm, _dev(x, cuda), _dev(q, cuda),
                        dv.element_tensor(el, cuda), v=_dev(v, cuda), precision=prec)
    assert np.linalg.norm(rho_o) > 0
    assert rel_l2(rho.cpu().numpy(), rho_o) < tol
    assert rel_l2(j.cpu().numpy(), j_o) < tol


def test_deposit_normalized_orders_match_oracle(order_case, cuda):
    """At R = 0.25 every ball holds several nodes even of P1 elements.  With R
    below the node spacing a ball can just touch one node: the integral I is
    then tiny, and q / I carries the relative error of t = R - |n - x|^2 / R
    near the edge of the support (the expanded form's ~1e-16 h^2 / R over t),
    7e-12 on this P1 mesh at R = 0.08 although I itself stays within 1e-12 of
    its bound."""
    from pyrticle_amd import device as dv
    wl, dm, om, x, q, v, el = order_case
    rho, j, rho_o, j_o, nz = _normalized_case(dm, om, wl, x, q, v, el, 0.25, 3, dv.FP64, cuda)
    assert nz == 0
    assert rel_l2(rho, rho_o) < FP64_TOL and rel_l2(j, j_o) < FP64_TOL


def test_average_forces_orders_match_oracle(order_case, cuda):
    wl, dm, om, x, q, v, el = order_case
    _average_case(dm, om, wl, x, q, v, el, _poly_fields(wl.ma.nodes, 3), wl.radius, 3, cuda,
                  vis=False)
