"""Pins the FP64 CPU gradient reference (oracle/adjoint_ref.py) on the CPU:
its forward equals the oracle's restatement of the reference (deposition and
the LSRK4 push), and its autograd gradients equal central finite differences.
"""

import numpy as np
import torch

from oracle import adjoint_ref as AR
from oracle import oracle as O
from tests.conftest import rel_l2


def _sample(wl, n, seed):
    rng = np.random.default_rng(seed)
    idx = rng.choice(wl.x.shape[0], size=n, replace=False)
    return (wl.x[idx].copy(), wl.p[idx].copy(), wl.q[idx].copy(), wl.m[idx].copy())


def test_deposit_matches_oracle_2d(square, oracle_square):
    x, p, q, m = _sample(square, 400, 1)
    v = O.velocities(p, m, square.c)
    el = np.array([O.find_containing_element(oracle_square, xi) for xi in x], dtype=np.uint32)
    rho_o, j_o, _ = O.deposit(oracle_square, square.radius, square.alpha, x, q, v, el)
    N = O.shape_normalizer(square.radius, 2, square.alpha)
    pairs = AR.NodePairs(square.ma.nodes, square.radius)
    rho, j = AR.deposit(pairs, torch.from_numpy(x), torch.from_numpy(q), torch.from_numpy(v),
                        square.alpha, N)
    assert rel_l2(rho.numpy(), rho_o) < 1e-13
    assert rel_l2(j.numpy(), j_o) < 1e-13


def test_push_matches_oracle_2d(square, oracle_square):
    """Uniform TE fields (E_y, B_z): the field function is the DG interpolant."""
    x, p, q, m = _sample(square, 300, 2)
    el = np.array([O.find_containing_element(oracle_square, xi) for xi in x], dtype=np.uint32)
    xo, po = x.copy(), p.copy()
    n, _ = O.lsrk4_run(oracle_square, square.fields, square.c, xo, po, q.copy(), m.copy(),
                       el.copy(), square.dt, 10)
    assert n == x.shape[0]

    def fields(xx):
        k = xx.shape[0]
        z = torch.zeros(k, dtype=torch.float64)
        return (torch.stack([z, torch.full_like(z, 1e3), z], dim=1),
                torch.stack([z, z, torch.full_like(z, 0.01)], dim=1))

    xt, pt = AR.push(torch.from_numpy(x), torch.from_numpy(p), torch.from_numpy(q),
                     torch.from_numpy(m), square.c, square.dt, 10, fields)
    assert rel_l2(xt.numpy(), xo) < 1e-14
    assert rel_l2(pt.numpy(), po) < 1e-13


def test_lsrk4_coefficients_match_oracle():
    a, b = O.lsrk4_coefficients()
    assert np.array_equal(a, np.array(AR.RK4A)) and np.array_equal(b, np.array(AR.RK4B))


def test_autograd_matches_finite_differences(square):
    """d/dx, d/dp of a linear functional of rho after a short push."""
    x, p, q, m = _sample(square, 60, 3)
    N = O.shape_normalizer(square.radius, 2, square.alpha)
    pairs = AR.NodePairs(square.ma.nodes, square.radius)
    rng = np.random.default_rng(4)
    w = torch.from_numpy(rng.normal(size=square.ma.nodes.shape[0]))

    def fields(xx):
        # non-uniform, to exercise the position dependence of the force
        X, Y = xx[:, 0], xx[:, 1]
        z = torch.zeros_like(X)
        return (torch.stack([1e3 * Y, 1e3 + 5e2 * X * X, z], dim=1),
                torch.stack([z, z, 0.01 + 0.005 * X * Y], dim=1))

    def loss(xx, pp):
        xT, pT = AR.push(xx, pp, torch.from_numpy(q), torch.from_numpy(m), square.c,
                         square.dt, 5, fields)
        rho, _ = AR.deposit(pairs, xT, torch.from_numpy(q), AR.velocities(
            pT, torch.from_numpy(m), square.c), square.alpha, N, want_j=False)
        return (rho * w).sum()

    xt = torch.from_numpy(x).requires_grad_(True)
    pt = torch.from_numpy(p).requires_grad_(True)
    loss(xt, pt).backward()
    dx = rng.normal(size=x.shape) * 1e-2
    dp = rng.normal(size=p.shape) * 1e-2 * np.abs(p).max()
    h = 1e-4
    with torch.no_grad():
        lp = float(loss(torch.from_numpy(x + h * dx), torch.from_numpy(p + h * dp)))
        lm = float(loss(torch.from_numpy(x - h * dx), torch.from_numpy(p - h * dp)))
    fd = (lp - lm) / (2 * h)
    ad = float((xt.grad.numpy() * dx).sum() + (pt.grad.numpy() * dp).sum())
    assert abs(fd - ad) <= 1e-6 * abs(ad)
