"""Parity of every C-ABI entry point with the oracle, on the GPU.

Integer work (element ids, kill order, sort permutation, counters) must match
bit for bit; floating-point results within the stated tolerances:
FP64 rho, j, forces and trajectories within 1e-12 relative L2; FP32C rho and j
within 1e-6 relative L2.
"""

import numpy as np
import pytest
import torch

from oracle import oracle as O
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
FP32C_TOL = 1e-6


def _dev(a, cuda, dtype=torch.float64):
    return torch.from_numpy(np.ascontiguousarray(a)).to(device=cuda, dtype=dtype)


@pytest.fixture(scope="module")
def dmesh2(square, cuda):
    from pyrticle_amd.device import DeviceMesh
    return DeviceMesh(square.ma)


@pytest.fixture(scope="module")
def dmesh3(small_cylinder, cuda):
    from pyrticle_amd.device import DeviceMesh
    return DeviceMesh(small_cylinder.ma)


def _points_2d(n, seed):
    rng = np.random.default_rng(seed)
    return rng.uniform(-1.05, 1.05, size=(n, 2))


def test_find_containing_element_bit_exact(square, oracle_square, dmesh2, cuda):
    from pyrticle_amd.device import element_numpy
    x = _points_2d(4000, 11)
    # add points exactly on vertices and edge midpoints (ties between elements)
    verts = square.ma.vertices[:200]
    mids = 0.5 * (square.ma.vertices[square.ma.el_vertices[:200, 0]]
                  + square.ma.vertices[square.ma.el_vertices[:200, 1]])
    x = np.vstack([x, verts, mids])
    got = element_numpy(dmesh2.find_containing_element(_dev(x, cuda)))
    want = np.array([O.find_containing_element(oracle_square, xi) for xi in x], dtype=np.uint32)
    assert np.array_equal(got, want)
    assert (want == O.INVALID).sum() > 0      # some points are outside the square


def _moved_particles(wl, om, scale, seed):
    rng = np.random.default_rng(seed)
    n = 3000
    idx = rng.choice(wl.x.shape[0], size=n, replace=False)
    x0 = wl.x[idx].copy()
    el0 = np.array([O.find_containing_element(om, xi) for xi in x0], dtype=np.uint32)
    x1 = x0 + rng.normal(0, scale, size=x0.shape)
    p = wl.p[idx].copy()
    return x1, p, wl.q[idx].copy(), wl.m[idx].copy(), el0


@pytest.mark.parametrize("scale", [0.002, 0.02, 0.2])
def test_locate_and_kill_order_bit_exact(square, oracle_square, dmesh2, cuda, scale):
    from pyrticle_amd import device as dv
    x, p, q, m, el = _moved_particles(square, oracle_square, scale, 5)
    ids = np.arange(x.shape[0], dtype=np.int64)
    # oracle: sequential update_containing_elements with kills
    ox, op, oq, om_, oel, oids = x.copy(), p.copy(), q.copy(), m.copy(), el.copy(), ids.copy()
    n_o, cnt_o = O.update_containing_elements(oracle_square, ox, op, oq, om_, oel, ids=oids)
    # GPU: locate, then replay the kill order
    tx, tp, tq, tm = (_dev(a, cuda) for a in (x, p, q, m))
    tel = dv.element_tensor(el, cuda)
    tids = _dev(ids, cuda, torch.int64)
    counters = torch.zeros(4, dtype=torch.int64, device=cuda)
    dead = dmesh2.locate(tx, tp, tel, counters)
    n_new, frm, to = dv.compact_plan(dead)
    for buf in (tx, tp, tq, tm, tel, tids):
        dv.apply_moves(buf, frm, to)
    torch.cuda.synchronize()
    assert n_new == n_o
    assert np.array_equal(dv.element_numpy(tel[:n_new]), oel[:n_o])
    assert np.array_equal(tids[:n_new].cpu().numpy(), oids[:n_o])
    assert np.array_equal(tx[:n_new].cpu().numpy(), ox[:n_o])
    assert tuple(int(c) for c in counters.cpu()) == cnt_o
    if scale == 0.2:
        assert n_new < x.shape[0]          # kills were exercised


def test_locate_periodic_wrap(periodic_cylinder, cuda):
    from pyrticle_amd import device as dv
    discr, ma = periodic_cylinder
    om = O.OracleMesh(ma)
    dm = dv.DeviceMesh(ma)
    rng = np.random.default_rng(3)
    n = 2000
    r = 0.8 * np.sqrt(rng.uniform(size=n))
    ph = rng.uniform(0, 2 * np.pi, n)
    x = np.stack([r * np.cos(ph), r * np.sin(ph), rng.uniform(0.01, 1.99, n)], axis=1)
    el = np.array([O.find_containing_element(om, xi) for xi in x], dtype=np.uint32)
    p = rng.normal(size=(n, 3))
    x1 = x.copy()
    x1[:, 2] += rng.normal(0, 0.5, n)           # many cross z=0 or z=2
    q = np.ones(n)
    m = np.ones(n)
    ox, oel = x1.copy(), el.copy()
    oids = np.arange(n, dtype=np.int64)
    n_o, cnt_o = O.update_containing_elements(om, ox, p.copy(), q.copy(), m.copy(), oel,
                                              ids=oids)
    tx = _dev(x1, cuda)
    tel = dv.element_tensor(el, cuda)
    tids = torch.arange(n, dtype=torch.int64, device=cuda)
    counters = torch.zeros(4, dtype=torch.int64, device=cuda)
    dead = dm.locate(tx, _dev(p, cuda), tel, counters)
    n_new, frm, to = dv.compact_plan(dead)
    for buf in (tx, tel, tids):
        dv.apply_moves(buf, frm, to)
    assert n_new == n_o
    assert np.array_equal(tids[:n_new].cpu().numpy(), oids[:n_o])
    assert np.array_equal(dv.element_numpy(tel[:n_new]), oel[:n_o])
    assert np.array_equal(tx[:n_new].cpu().numpy(), ox[:n_o])
    assert tuple(int(c) for c in counters.cpu()) == cnt_o


def test_velocities(square, cuda):
    from pyrticle_amd import device as dv
    p = square.p.copy()
    p[:5] = 0.0                                  # the |p| == 0 branch
    v = dv.velocities(_dev(p, cuda), _dev(square.m, cuda), square.c).cpu().numpy()
    vo = O.velocities(p, square.m, square.c)
    assert np.max(np.abs(v - vo) / np.maximum(np.abs(vo), 1e-300)) <= 4.5e-16
    assert np.all(v[:5] == 0)


def test_forces_2d(square, oracle_square, dmesh2, cuda):
    from pyrticle_amd import device as dv
    n = 4000
    x, p, q, m = square.x[:n], square.p[:n], square.q[:n], square.m[:n]
    el = np.array([O.find_containing_element(oracle_square, xi) for xi in x], dtype=np.uint32)
    v = O.velocities(p, m, square.c)
    rng = np.random.default_rng(2)
    nn = square.ma.nnodes
    fields = [rng.normal(size=nn), rng.normal(size=nn), None, None, None, rng.normal(size=nn)]
    fo, viso = O.forces(oracle_square, x, el, q, v, fields, vis=True)
    tf = [None if f is None else _dev(f, cuda) for f in fields]
    F, vis = dmesh2.forces(tf, _dev(x, cuda), dv.element_tensor(el, cuda), _dev(q, cuda),
                           _dev(v, cuda), vis=True)
    assert rel_l2(F.cpu().numpy(), fo) < 1e-13
    for k in viso:
        assert rel_l2(vis[k].cpu().numpy(), viso[k]) < 1e-13, k


def test_forces_3d(small_cylinder, cuda, dmesh3):
    from pyrticle_amd import device as dv
    wl = small_cylinder
    om = O.OracleMesh(wl.ma)
    n = 2000
    x, p, q, m = wl.x[:n], wl.p[:n], wl.q[:n], wl.m[:n]
    el = dv.element_numpy(dmesh3.find_containing_element(_dev(x, cuda)))
    v = O.velocities(p, m, wl.c)
    rng = np.random.default_rng(4)
    fields = [rng.normal(size=wl.ma.nnodes) for _ in range(6)]
    fo = O.forces(om, x, el, q, v, fields)
    F = dmesh3.forces([_dev(f, cuda) for f in fields], _dev(x, cuda),
                      dv.element_tensor(el, cuda), _dev(q, cuda), _dev(v, cuda))
    assert rel_l2(F.cpu().numpy(), fo) < 1e-13


@pytest.mark.parametrize("n", [50_000, 1_500_000])
def test_sort_by_element_is_stable_argsort(square, dmesh2, cuda, n):
    """Both sort paths (LSD passes below 2^20 keys, rocPRIM's onesweep sort
    above), with 1% of the particles outside the mesh (their ids clamp to K and
    sort last, in index order)."""
    from pyrticle_amd import device as dv
    rng = np.random.default_rng(9)
    el = rng.integers(0, square.ma.K, size=n).astype(np.uint32)
    el[rng.random(n) < 0.01] = dv.INVALID_ELEMENT
    perm, seg = dmesh2.sort_by_element(dv.element_tensor(el, cuda))
    assert np.array_equal(perm.cpu().numpy(), np.argsort(el, kind="stable"))
    counts = np.bincount(el[el != dv.INVALID_ELEMENT], minlength=square.ma.K)
    assert np.array_equal(seg.cpu().numpy(), np.concatenate([[0], np.cumsum(counts)]))


@pytest.mark.parametrize("precision", ["fp64", "fp32c"])
def test_deposit_2d_matches_oracle(square, oracle_square, dmesh2, cuda, precision):
    from pyrticle_amd import device as dv
    wl = square
    el = np.array([O.find_containing_element(oracle_square, xi) for xi in wl.x],
                  dtype=np.uint32)
    v = O.velocities(wl.p, wl.m, wl.c)
    rho_o, j_o, visits = O.deposit(oracle_square, wl.radius, wl.alpha, wl.x, wl.q, v, el)
    nrm = O.shape_normalizer(wl.radius, 2, wl.alpha)
    prec = dv.FP64 if precision == "fp64" else dv.FP32C
    rho, j = dmesh2.deposit(wl.radius, wl.alpha, nrm, _dev(wl.x, cuda), _dev(wl.q, cuda),
                            dv.element_tensor(el, cuda), v=_dev(v, cuda), precision=prec)
    tol = FP64_TOL if precision == "fp64" else FP32C_TOL
    assert rel_l2(rho.cpu().numpy(), rho_o) < tol
    assert rel_l2(j.cpu().numpy(), j_o) < tol


@pytest.mark.parametrize("precision", ["fp64", "fp32c"])
def test_deposit_3d_matches_oracle(small_cylinder, dmesh3, cuda, precision):
    from pyrticle_amd import device as dv
    wl = small_cylinder
    om = O.OracleMesh(wl.ma)
    n = 5000
    x, q = wl.x[:n], wl.q[:n]
    el = dv.element_numpy(dmesh3.find_containing_element(_dev(x, cuda)))
    v = O.velocities(wl.p[:n], wl.m[:n], wl.c)
    rho_o, j_o, _ = O.deposit(om, wl.radius, wl.alpha, x, q, v, el)
    nrm = O.shape_normalizer(wl.radius, 3, wl.alpha)
    prec = dv.FP64 if precision == "fp64" else dv.FP32C
    tol = FP64_TOL if precision == "fp64" else FP32C_TOL
    rho, j = dmesh3.deposit(wl.radius, wl.alpha, nrm, _dev(x, cuda), _dev(q, cuda),
                            dv.element_tensor(el, cuda), v=_dev(v, cuda), precision=prec)
    assert rel_l2(rho.cpu().numpy(), rho_o) < tol
    assert rel_l2(j.cpu().numpy(), j_o) < tol
    rho2, j2 = dmesh3.deposit(wl.radius, wl.alpha, nrm, _dev(x, cuda), _dev(q, cuda),
                              dv.element_tensor(el, cuda), v=_dev(v, cuda), precision=prec)
    assert torch.equal(rho, rho2) and torch.equal(j, j2)     # deterministic


def test_deposit_counting_mode(small_cylinder, dmesh3, cuda):
    """The pair-counting mode runs the candidate gather (every candidate
    particle of every target, the target test, all nodes of the target): its
    rho and j equal the list gather's to rounding, it counts exactly the
    node-particle pairs closer than R (a k-d tree brute force) and evaluates
    every node of every particle that passes the target test."""
    from scipy.spatial import cKDTree
    from pyrticle_amd import device as dv
    wl = small_cylinder
    n = 5000
    x, q = wl.x[:n], wl.q[:n]
    el = dmesh3.find_containing_element(_dev(x, cuda))
    v = O.velocities(wl.p[:n], wl.m[:n], wl.c)
    nrm = O.shape_normalizer(wl.radius, 3, wl.alpha)
    args = (wl.radius, wl.alpha, nrm, _dev(x, cuda), _dev(q, cuda), el)
    rho_l, j_l = dmesh3.deposit(*args, v=_dev(v, cuda))
    dmesh3.set_deposit_counting(True)
    try:
        rho_c, j_c = dmesh3.deposit(*args, v=_dev(v, cuda))
        n_eval, n_in, n_cand, n_pass = dmesh3.deposit_stats()
    finally:
        dmesh3.set_deposit_counting(False)
    assert rel_l2(rho_c.cpu().numpy(), rho_l.cpu().numpy()) < 1e-13
    assert rel_l2(j_c.cpu().numpy(), j_l.cpu().numpy()) < 1e-13
    tree = cKDTree(np.asarray(wl.ma.nodes).reshape(-1, 3))
    inside = sum(len(h) for h in tree.query_ball_point(x, wl.radius * (1 - 1e-12)))
    touching = sum(len(h) for h in tree.query_ball_point(x, wl.radius * (1 + 1e-12)))
    assert inside <= n_in <= touching
    assert n_pass <= n_cand and n_eval == n_pass * wl.ma.Np
    with pytest.raises(ValueError):   # FP32C has no counting pass
        dmesh3.set_deposit_counting(True)
        try:
            dmesh3.deposit(*args, v=_dev(v, cuda), precision=dv.FP32C)
        finally:
            dmesh3.set_deposit_counting(False)


@pytest.mark.parametrize("precision", ["fp64", "fp32c"])
def test_deposit_dense_beam_matches_oracle(small_cylinder, dmesh3, cuda, precision):
    """Thousands of particles per element: the gather's LDS queue runs deep."""
    from pyrticle_amd import device as dv
    from pyrticle_amd.workloads import cylinder_beam
    wl = small_cylinder
    om = O.OracleMesh(wl.ma)
    x, p, q, m = cylinder_beam(120_000, seed=5, axes=(0.08, 0.06), sigma_z=0.05)
    el = dv.element_numpy(dmesh3.find_containing_element(_dev(x, cuda)))
    assert np.bincount(el).max() > 2000
    v = O.velocities(p, m, wl.c)
    rho_o, j_o, _ = O.deposit(om, wl.radius, wl.alpha, x, q, v, el)
    nrm = O.shape_normalizer(wl.radius, 3, wl.alpha)
    prec = dv.FP64 if precision == "fp64" else dv.FP32C
    rho, j = dmesh3.deposit(wl.radius, wl.alpha, nrm, _dev(x, cuda), _dev(q, cuda),
                            dv.element_tensor(el, cuda), v=_dev(v, cuda), precision=prec)
    tol = FP64_TOL if precision == "fp64" else FP32C_TOL
    assert rel_l2(rho.cpu().numpy(), rho_o) < tol
    assert rel_l2(j.cpu().numpy(), j_o) < tol


def _normalized_case(dmesh, om, wl, x, q, v, el, radius, dim, prec, cuda):
    from pyrticle_amd import device as dv
    w = wl.discr.element_groups[0].local_discretization.integral_weights()
    rho_o, j_o, s_o, nz_o = O.deposit_normalized(om, radius, wl.alpha, x, q, v, el, w)
    nrm = O.shape_normalizer(radius, dim, wl.alpha)
    scale = torch.empty(x.shape[0], dtype=torch.float64, device=cuda)
    rho, j, nz = dmesh.deposit_normalized(radius, wl.alpha, nrm, _dev(x, cuda), _dev(q, cuda),
                                          dv.element_tensor(el, cuda), w, v=_dev(v, cuda),
                                          precision=prec, scale=scale)
    s = scale.cpu().numpy()
    assert nz == nz_o
    assert np.array_equal(s == 0, s_o == 0)
    live = s_o != 0
    assert rel_l2(s, s_o) < FP64_TOL
    # Per particle, the shape integral I = q / scale.  Each node term
    # |J| w S(n - x) carries an error of a few ulps of |J| |w| max(S) (R - r^2/R
    # cancels near the support's edge), so a support that barely reaches a node
    # makes I small and its relative error large: the bound is that sum of
    # |J| |w| max(S) over the nodes within R, not I itself.
    from scipy.spatial import cKDTree
    node_w = np.abs(wl.ma.jacobian[:, None] * w[None, :]).ravel()
    near = cKDTree(wl.ma.nodes).query_ball_point(x[live], radius)
    bound = np.array([node_w[i].sum() for i in near]) * nrm * radius ** wl.alpha
    integ, integ_o = q[live] / s[live], q[live] / s_o[live]
    assert (np.abs(integ - integ_o) <= FP64_TOL * bound).all()
    return rho.cpu().numpy(), j.cpu().numpy(), rho_o, j_o, nz


@pytest.mark.parametrize("precision", ["fp64", "fp32c"])
def test_deposit_normalized_2d_matches_oracle(square, oracle_square, dmesh2, cuda, precision):
    """NormalizedShapeFunctionDepositor's deposit: the beam plus particles whose
    shapes are cut off by the boundary."""
    from pyrticle_amd import device as dv
    wl = square
    rng = np.random.default_rng(11)
    x = np.concatenate([wl.x, rng.uniform(-0.99, 0.99, size=(3000, 2))])
    el = dv.element_numpy(dmesh2.find_containing_element(_dev(x, cuda)))
    n = x.shape[0]
    q = np.full(n, wl.q[0])
    v = O.velocities(np.resize(wl.p, (n, 2)), np.resize(wl.m, n), wl.c)
    prec = dv.FP64 if precision == "fp64" else dv.FP32C
    tol = FP64_TOL if precision == "fp64" else FP32C_TOL
    rho, j, rho_o, j_o, nz = _normalized_case(dmesh2, oracle_square, wl, x, q, v, el,
                                              wl.radius, 2, prec, cuda)
    assert nz == 0
    assert rel_l2(rho, rho_o) < tol and rel_l2(j, j_o) < tol
    assert abs(wl.discr.integral(rho) / q.sum() - 1) < (1e-12 if precision == "fp64" else 1e-6)


@pytest.mark.parametrize("precision", ["fp64", "fp32c"])
def test_deposit_normalized_3d_matches_oracle(small_cylinder, dmesh3, cuda, precision):
    from pyrticle_amd import device as dv
    wl = small_cylinder
    om = O.OracleMesh(wl.ma)
    n = 5000
    x, q = wl.x[:n], wl.q[:n]
    el = dv.element_numpy(dmesh3.find_containing_element(_dev(x, cuda)))
    v = O.velocities(wl.p[:n], wl.m[:n], wl.c)
    prec = dv.FP64 if precision == "fp64" else dv.FP32C
    tol = FP64_TOL if precision == "fp64" else FP32C_TOL
    rho, j, rho_o, j_o, _ = _normalized_case(dmesh3, om, wl, x, q, v, el, wl.radius, 3, prec,
                                             cuda)
    assert rel_l2(rho, rho_o) < tol and rel_l2(j, j_o) < tol


def test_deposit_normalized_zero_integral(square, oracle_square, dmesh2, cuda):
    """A support far below the node spacing: most particles reach no node, deposit
    nothing and are counted; the count and the zero pattern match the oracle."""
    from pyrticle_amd import device as dv
    wl = square
    rng = np.random.default_rng(2)
    x = np.concatenate([rng.uniform(-0.9, 0.9, size=(400, 2)),
                        wl.ma.nodes[[17, 403, 999, 4000]] + 1e-5])
    el = dv.element_numpy(dmesh2.find_containing_element(_dev(x, cuda)))
    q = np.full(x.shape[0], wl.q[0])
    v = np.zeros_like(x)
    rho, _, rho_o, _, nz = _normalized_case(dmesh2, oracle_square, wl, x, q, v, el, 1e-3, 2,
                                            dv.FP64, cuda)
    assert 300 < nz < x.shape[0] - 4
    assert rel_l2(rho, rho_o) < FP64_TOL


def _poly_fields(nodes, dim):
    x = nodes
    z = x[:, 2] if dim == 3 else 0.0
    return [1e3 * (1 + x[:, 0] ** 2), -2e3 * x[:, 1] * x[:, 0], 5e2 * (1 + z),
            0.3 * x[:, 1], 0.2 + 0.1 * z, 0.5 - 0.4 * x[:, 0] * x[:, 1]]


def _average_case(dmesh, om, wl, x, q, v, el, fields, radius, dim, cuda, vis):
    from pyrticle_amd import device as dv
    w = wl.discr.element_groups[0].local_discretization.integral_weights()
    ref = O.average_forces(om, radius, wl.alpha, x, q, v, el, w, fields, vis=vis)
    nrm = O.shape_normalizer(radius, dim, wl.alpha)
    dfields = [None if f is None else _dev(f, cuda) for f in fields]
    out = dmesh.average_forces(radius, wl.alpha, nrm, dfields, _dev(x, cuda),
                               dv.element_tensor(el, cuda), _dev(q, cuda), _dev(v, cuda), w,
                               vis=vis)
    f, integ = out[0].cpu().numpy(), out[1].cpu().numpy()
    assert out[2] == ref[2]
    assert np.array_equal(integ == 0, ref[1] == 0)
    assert rel_l2(q * integ, ref[1]) < FP64_TOL
    assert rel_l2(f, ref[0]) < FP64_TOL
    return out, ref


def test_average_forces_2d_matches_oracle(square, oracle_square, dmesh2, cuda):
    """Averaging pusher, TE fields (ex, ey, bz), polynomial in x, with particles
    whose shapes the boundary cuts: forces, averages, deviations."""
    from pyrticle_amd import device as dv
    wl = square
    rng = np.random.default_rng(4)
    x = np.concatenate([wl.x, rng.uniform(-0.98, 0.98, size=(3000, 2))])
    el = dv.element_numpy(dmesh2.find_containing_element(_dev(x, cuda)))
    n = x.shape[0]
    q = np.full(n, wl.q[0])
    v = O.velocities(np.resize(wl.p, (n, 2)), np.resize(wl.m, n), wl.c)
    fields = _poly_fields(wl.ma.nodes, 2)
    fields[2] = fields[3] = fields[4] = None
    out, ref = _average_case(dmesh2, oracle_square, wl, x, q, v, el, fields, wl.radius, 2, cuda,
                             vis=True)
    vis, vref = out[3], ref[3]
    for name in ("pt_e", "pt_b", "el_force", "mag_force"):
        assert rel_l2(vis[name].cpu().numpy(), vref[name]) < FP64_TOL, name
    # deviations: sqrt of a difference of nearly equal means, so an absolute
    # bound relative to the field scale
    for name, ref_scale in (("pt_e_stddev", vref["pt_e"]), ("pt_b_stddev", vref["pt_b"])):
        err = np.abs(vis[name].cpu().numpy() - vref[name]).max()
        assert err < 1e-6 * np.abs(ref_scale).max(), name


@pytest.mark.parametrize("radius,alpha", [(0.6, 2.0), (0.15, 3.0), (0.15, 1.5)])
def test_shape_radius_and_exponent(square, oracle_square, dmesh2, cuda, radius, alpha):
    """Shapes beside the FP64 kernels' alpha = 2, R <= 1/2 form (one clamped
    multiply per node): R > 1/2 squares explicitly, alpha != 2 calls pow.
    Plain deposit (FP64, and FP32C, whose general-alpha path uses the FP32
    pow: within 1e-5), normalized deposit and averaging forces (FP64)."""
    from pyrticle_amd import device as dv
    wl = square
    n = 1500
    x, q = wl.x[:n], wl.q[:n]
    el = np.array([O.find_containing_element(oracle_square, xi) for xi in x], dtype=np.uint32)
    v = O.velocities(wl.p[:n], wl.m[:n], wl.c)
    nrm = O.shape_normalizer(radius, 2, alpha)
    rho_o, j_o, _ = O.deposit(oracle_square, radius, alpha, x, q, v, el)
    for prec, tol in ((dv.FP64, FP64_TOL), (dv.FP32C, 1e-6 if alpha == 2.0 else 1e-5)):
        rho, j = dmesh2.deposit(radius, alpha, nrm, _dev(x, cuda), _dev(q, cuda),
                                dv.element_tensor(el, cuda), v=_dev(v, cuda), precision=prec)
        assert rel_l2(rho.cpu().numpy(), rho_o) < tol, prec
        assert rel_l2(j.cpu().numpy(), j_o) < tol, prec
    w = wl.discr.element_groups[0].local_discretization.integral_weights()
    rho_o, j_o, s_o, nz_o = O.deposit_normalized(oracle_square, radius, alpha, x, q, v, el, w)
    rho, j, nz = dmesh2.deposit_normalized(radius, alpha, nrm, _dev(x, cuda), _dev(q, cuda),
                                           dv.element_tensor(el, cuda), w, v=_dev(v, cuda),
                                           precision=dv.FP64)
    assert nz == nz_o
    assert rel_l2(rho.cpu().numpy(), rho_o) < FP64_TOL
    assert rel_l2(j.cpu().numpy(), j_o) < FP64_TOL
    fields = _poly_fields(wl.ma.nodes, 2)
    fields[2] = fields[3] = fields[4] = None
    ref = O.average_forces(oracle_square, radius, alpha, x, q, v, el, w, fields, vis=False)
    out = dmesh2.average_forces(radius, alpha, nrm, [None if f is None else _dev(f, cuda)
                                                     for f in fields],
                                _dev(x, cuda), dv.element_tensor(el, cuda), _dev(q, cuda),
                                _dev(v, cuda), w, vis=False)
    assert out[2] == ref[2]
    assert rel_l2(q * out[1].cpu().numpy(), ref[1]) < FP64_TOL
    assert rel_l2(out[0].cpu().numpy(), ref[0]) < FP64_TOL


def test_average_forces_3d_matches_oracle(small_cylinder, dmesh3, cuda):
    from pyrticle_amd import device as dv
    wl = small_cylinder
    om = O.OracleMesh(wl.ma)
    n = 4000
    x, q = wl.x[:n], wl.q[:n]
    el = dv.element_numpy(dmesh3.find_containing_element(_dev(x, cuda)))
    v = O.velocities(wl.p[:n], wl.m[:n], wl.c)
    _average_case(dmesh3, om, wl, x, q, v, el, _poly_fields(wl.ma.nodes, 3), wl.radius, 3, cuda,
                  vis=False)


def test_average_forces_zero_charge(square, oracle_square, dmesh2, cuda):
    from pyrticle_amd import device as dv
    wl = square
    rng = np.random.default_rng(2)
    x = np.concatenate([rng.uniform(-0.9, 0.9, size=(400, 2)),
                        wl.ma.nodes[[17, 403, 999]] + 1e-5])
    el = dv.element_numpy(dmesh2.find_containing_element(_dev(x, cuda)))
    q = np.full(x.shape[0], wl.q[0])
    v = np.full_like(x, 1e6)
    out, ref = _average_case(dmesh2, oracle_square, wl, x, q, v, el, wl.fields, 1e-3, 2, cuda,
                             vis=True)
    assert 300 < out[2] < x.shape[0] - 3
    dead = (ref[1] == 0)
    for name, arr in out[3].items():
        assert not arr.cpu().numpy()[dead].any(), name


@pytest.fixture(scope="module")
def periodic_beam(periodic_cylinder):
    """Particles in the periodic cylinder (z period 2), half of them within R of
    the periodic faces, so that their shapes wrap around."""
    from pyrticle_amd.units import SIUnitsWithNaturalConstants as U
    discr, ma = periodic_cylinder
    om = O.OracleMesh(ma)
    rng = np.random.default_rng(21)
    n = 1500
    r = 0.7 * np.sqrt(rng.uniform(size=n))
    ph = rng.uniform(0, 2 * np.pi, n)
    z = np.where(rng.uniform(size=n) < 0.5, rng.uniform(0.001, 0.25, n),
                 rng.uniform(1.75, 1.999, n))
    x = np.stack([r * np.cos(ph), r * np.sin(ph), z], axis=1)
    el = np.array([O.find_containing_element(om, xi) for xi in x], dtype=np.uint32)
    assert (el != 0xFFFFFFFF).all()
    c = U.VACUUM_LIGHT_SPEED()
    v = rng.normal(0, 0.1 * c, size=(n, 3))
    q = np.full(n, -U.EL_CHARGE)
    return discr, ma, om, x, el, v, q


@pytest.mark.parametrize("precision", ["fp64", "fp32c"])
def test_deposits_periodic_match_oracle(periodic_beam, cuda, precision):
    """Plain and normalized deposition with shapes that wrap through the
    periodic faces (shift codes on the candidate lists)."""
    from pyrticle_amd import device as dv
    discr, ma, om, x, el, v, q = periodic_beam
    dm = dv.DeviceMesh(ma)
    R, alpha = 0.3, 2.0
    nrm = O.shape_normalizer(R, 3, alpha)
    prec = dv.FP64 if precision == "fp64" else dv.FP32C
    tol = FP64_TOL if precision == "fp64" else FP32C_TOL
    rho_o, j_o, _ = O.deposit(om, R, alpha, x, q, v, el)
    rho, j = dm.deposit(R, alpha, nrm, _dev(x, cuda), _dev(q, cuda), dv.element_tensor(el, cuda),
                        v=_dev(v, cuda), precision=prec)
    assert rel_l2(rho.cpu().numpy(), rho_o) < tol and rel_l2(j.cpu().numpy(), j_o) < tol
    w = discr.element_groups[0].local_discretization.integral_weights()
    rho_o, j_o, s_o, nz_o = O.deposit_normalized(om, R, alpha, x, q, v, el, w)
    scale = torch.empty(x.shape[0], dtype=torch.float64, device=cuda)
    rho, j, nz = dm.deposit_normalized(R, alpha, nrm, _dev(x, cuda), _dev(q, cuda),
                                       dv.element_tensor(el, cuda), w, v=_dev(v, cuda),
                                       precision=prec, scale=scale)
    assert nz == nz_o == 0
    assert rel_l2(scale.cpu().numpy(), s_o) < FP64_TOL
    assert rel_l2(rho.cpu().numpy(), rho_o) < tol and rel_l2(j.cpu().numpy(), j_o) < tol
    # the periodic mesh has no boundary: the normalized deposit holds the charge
    assert abs(discr.integral(rho.cpu().numpy()) / q.sum() - 1) < (1e-12 if precision == "fp64"
                                                                   else 1e-6)


def test_average_forces_periodic_match_oracle(periodic_beam, cuda):
    from pyrticle_amd import device as dv
    discr, ma, om, x, el, v, q = periodic_beam
    dm = dv.DeviceMesh(ma)
    R, alpha = 0.3, 2.0
    w = discr.element_groups[0].local_discretization.integral_weights()
    nodes = ma.nodes
    fields = [1e3 * (1 + nodes[:, 0]), 2e2 * nodes[:, 1] * nodes[:, 0], 5e2 * np.sin(np.pi * nodes[:, 2]),
              0.1 * nodes[:, 1], 0.2 + 0.05 * np.cos(np.pi * nodes[:, 2]), 0.5 - 0.3 * nodes[:, 0]]
    f_o, charge_o, nz_o, vis_o = O.average_forces(om, R, alpha, x, q, v, el, w, fields, vis=True)
    out = dm.average_forces(R, alpha, O.shape_normalizer(R, 3, alpha),
                            [_dev(f, cuda) for f in fields], _dev(x, cuda),
                            dv.element_tensor(el, cuda), _dev(q, cuda), _dev(v, cuda), w, vis=True)
    assert out[2] == nz_o == 0
    assert rel_l2(out[0].cpu().numpy(), f_o) < FP64_TOL
    assert rel_l2(q * out[1].cpu().numpy(), charge_o) < FP64_TOL
    for name in ("pt_e", "pt_b", "el_force", "mag_force"):
        assert rel_l2(out[3][name].cpu().numpy(), vis_o[name]) < FP64_TOL, name


@pytest.mark.parametrize("case", ["square", "cylinder", "periodic"])
def test_face_walk_stats_match_oracle(case, square, oracle_square, small_cylinder,
                                      periodic_beam, dmesh2, dmesh3, cuda):
    """The reference's face walk for the normalized depositor's statistics:
    elements per particle bit-exact, centroid-distance sums, squares and
    extremes within 1e-13 (periodic images included)."""
    from pyrticle_amd import device as dv
    if case == "square":
        om, dm, x, R = oracle_square, dmesh2, square.x, square.radius
        el = np.array([O.find_containing_element(om, xi) for xi in x], dtype=np.uint32)
    elif case == "cylinder":
        wl = small_cylinder
        om, dm, x, R = O.OracleMesh(wl.ma), dmesh3, wl.x[:3000], wl.radius
        el = dv.element_numpy(dm.find_containing_element(_dev(x, cuda)))
    else:
        _, ma, om, x, el, _, _ = periodic_beam
        dm, R = dv.DeviceMesh(ma), 0.3
    ref = O.face_walk_stats(om, R, x, el)
    got = [t.cpu().numpy() for t in dm.face_walk_stats(R, _dev(x, cuda), dv.element_tensor(el, cuda))]
    assert np.array_equal(got[0].astype(np.uint32), ref[0])
    for k in range(1, 5):
        assert np.allclose(got[k], ref[k], rtol=1e-13, atol=0), k


def test_new_entry_points_edge_cases(square, oracle_square, dmesh2, cuda):
    """Empty inputs and particles without an element: zero outputs, no counts;
    a radius that breaks the face walk's bound fails loudly."""
    from pyrticle_amd import device as dv
    wl = square
    w = wl.discr.element_groups[0].local_discretization.integral_weights()
    nrm = O.shape_normalizer(wl.radius, 2, wl.alpha)
    x0 = torch.zeros((0, 2), dtype=torch.float64, device=cuda)
    q0 = torch.zeros(0, dtype=torch.float64, device=cuda)
    e0 = dv.element_tensor(np.zeros(0, dtype=np.uint32), cuda)
    rho, j, nz = dmesh2.deposit_normalized(wl.radius, wl.alpha, nrm, x0, q0, e0, w, v=x0)
    assert nz == 0 and not rho.abs().max() > 0 and not j.abs().max() > 0
    f, integ, nz = dmesh2.average_forces(wl.radius, wl.alpha, nrm, [None] * 6, x0, e0, q0, x0, w)
    assert f.shape == (0, 2) and nz == 0
    assert dmesh2.face_walk_stats(wl.radius, x0, e0)[0].numel() == 0
    # one located particle and one outside the mesh (element id INVALID)
    x = np.array([[0.1, 0.2], [5.0, 5.0]])
    el = np.array([O.find_containing_element(oracle_square, x[0]), 0xFFFFFFFF], dtype=np.uint32)
    fields = [_dev(f_, cuda) if f_ is not None else None for f_ in wl.fields]
    f, integ, nz = dmesh2.average_forces(wl.radius, wl.alpha, nrm, fields, _dev(x, cuda),
                                         dv.element_tensor(el, cuda),
                                         _dev(np.full(2, wl.q[0]), cuda),
                                         _dev(np.full((2, 2), 1e6), cuda), w)
    f = f.cpu().numpy()
    assert nz == 0 and np.abs(f[0]).max() > 0 and not f[1].any()
    assert integ.cpu().numpy()[1] == 0
    n_el = dmesh2.face_walk_stats(wl.radius, _dev(x, cuda), dv.element_tensor(el, cuda))[0]
    assert n_el.cpu().numpy()[1] == 0 and n_el.cpu().numpy()[0] > 0
    with pytest.raises(ValueError, match="face walk"):
        dmesh2.face_walk_stats(3.0, _dev(x[:1], cuda), dv.element_tensor(el[:1], cuda))


@pytest.mark.parametrize("precision", ["fp64", "fp32c"])
def test_deposit_list_edge_cases(small_cylinder, dmesh3, cuda, precision):
    """Inputs that shape the chunk lists unusually, against the oracle: no
    particles, only particles without an element, a single particle, many
    particles at one point (full chunks of identical positions), and one
    particle in each of many elements (chunks of one)."""
    from pyrticle_amd import device as dv
    wl = small_cylinder
    om = O.OracleMesh(wl.ma)
    nrm = O.shape_normalizer(wl.radius, 3, wl.alpha)
    prec = dv.FP64 if precision == "fp64" else dv.FP32C
    tol = FP64_TOL if precision == "fp64" else FP32C_TOL

    def run(x, el, v, q):
        rho, j = dmesh3.deposit(wl.radius, wl.alpha, nrm, _dev(x, cuda), _dev(q, cuda),
                                dv.element_tensor(el, cuda), v=_dev(v, cuda), precision=prec)
        return rho.cpu().numpy(), j.cpu().numpy()

    # no particles, and particles none of which has an element: zeros
    for n in (0, 5):
        x = np.tile(wl.x[:1], (n, 1))
        rho, j = run(x, np.full(n, 0xFFFFFFFF, dtype=np.uint32), np.ones((n, 3)),
                     np.full(n, wl.q[0]))
        assert not np.abs(rho).max(initial=0) and not np.abs(j).max(initial=0)
    el_all = dv.element_numpy(dmesh3.find_containing_element(_dev(wl.x, cuda)))
    v_all = O.velocities(wl.p, wl.m, wl.c)
    cases = {
        "single": np.array([0]),
        "coincident": np.zeros(300, dtype=np.int64),
        "one_per_element": np.unique(el_all, return_index=True)[1][:2000],
    }
    for name, idx in cases.items():
        x, el, v, q = wl.x[idx], el_all[idx], v_all[idx], wl.q[idx]
        rho_o, j_o, _ = O.deposit(om, wl.radius, wl.alpha, x, q, v, el)
        rho, j = run(x, el, v, q)
        assert rel_l2(rho, rho_o) < tol and rel_l2(j, j_o) < tol, name


def test_diagnostics_match_oracle(cuda):
    """The seven diagnostics on 1M particles against the oracle's sequential
    sums (summation order differs: sums within 1e-12, the cancelling emittance
    and one-pass energy spread within their conditioning)."""
    from pyrticle_amd import device as dv
    from tests.test_oracle import _diag_beam
    x, p, m, q, c = _diag_beam(n=1_000_000, seed=9)
    v = O.velocities(p, m, c)
    X, P, M, Q, V = (_dev(a, cuda) for a in (x, p, m, q, v))
    e = dv.kinetic_energies(P, M, c).cpu().numpy()
    e_o = O.kinetic_energies(p, m, c)
    assert np.abs(e / e_o - 1).max() < 1e-13
    assert dv.total_charge(Q) == pytest.approx(O.total_charge(q), rel=1e-12)
    assert np.allclose(dv.particle_momentum(P), O.particle_momentum(p), rtol=1e-12, atol=0)
    assert np.allclose(dv.particle_current(Q, V, 4.0), O.particle_current(q, v, 4.0),
                       rtol=1e-12, atol=0)
    for axis in range(3):
        assert dv.rms_beam_size(X, axis) == pytest.approx(O.rms_beam_size(x, axis), rel=1e-12)
    for axis in (0, 1):
        assert dv.rms_beam_emittance(X, P, axis, 2) == pytest.approx(
            O.rms_beam_emittance(x, p, axis, 2), rel=1e-9)
    cond = (e_o.mean() / e_o.std()) ** 2
    assert dv.rms_energy_spread(P, M, c) == pytest.approx(O.rms_energy_spread(p, m, c),
                                                          rel=1e-14 * cond)
    # deterministic, and an empty state gives zeros
    assert dv.rms_energy_spread(P, M, c) == dv.rms_energy_spread(P, M, c)
    assert dv.total_charge(Q[:0]) == 0 and dv.rms_beam_size(X[:0], 0) == 0


@pytest.mark.parametrize("dt_scale", [1.0, 4.0])
def test_lsrk4_square_matches_oracle(square, oracle_square, dmesh2, cuda, dt_scale):
    """Square workload, 10k electrons, 100 LSRK4 steps with relocalization; with a 4x
    time step particles leave the square and the kill order is exercised."""
    import dataclasses
    from pyrticle_amd import device as dv
    wl = dataclasses.replace(square, dt=square.dt * dt_scale)
    el = np.array([O.find_containing_element(oracle_square, xi) for xi in wl.x],
                  dtype=np.uint32)
    n0 = wl.x.shape[0]
    ox, op, oq, om_, oel = wl.x.copy(), wl.p.copy(), wl.q.copy(), wl.m.copy(), el.copy()
    oids = np.arange(n0, dtype=np.int64)
    n_o, cnt_o = O.lsrk4_run(oracle_square, wl.fields, wl.c, ox, op, oq, om_, oel, wl.dt,
                             wl.nsteps, ids=oids)
    tf = [None if f is None else _dev(f, cuda) for f in wl.fields]
    dmesh2.set_static_fields(tf)
    tx, tp, tq, tm = (_dev(a, cuda) for a in (wl.x, wl.p, wl.q, wl.m))
    tel = dv.element_tensor(el, cuda)
    tids = torch.arange(n0, dtype=torch.int64, device=cuda)
    counters = torch.zeros(4, dtype=torch.int64, device=cuda)
    n = dmesh2.lsrk4_steps(wl.nsteps, wl.dt, wl.c, n0, tx, tp, tq, tm, tel, ids=tids,
                           counters=counters)
    assert n == n_o
    if dt_scale > 1:
        assert n < n0                            # particles left the square
    assert np.array_equal(tids[:n].cpu().numpy(), oids[:n_o])
    assert np.array_equal(dv.element_numpy(tel[:n]), oel[:n_o])
    assert rel_l2(tx[:n].cpu().numpy(), ox[:n_o]) < FP64_TOL
    assert rel_l2(tp[:n].cpu().numpy(), op[:n_o]) < FP64_TOL
    assert tuple(int(c) for c in counters.cpu()) == cnt_o


@pytest.mark.parametrize("dt_scale", [1.0, 4.0])
def test_lsrk4_in_deposition_order_matches_oracle(square, oracle_square, cuda, dt_scale):
    """A deposit before every step, as in a PIC loop: the fused push then steps
    the particles in the order of the deposition's element sort (the square's
    particles are not element-sorted, so the order is a real shuffle; a step
    with kills drops it).  Ids, kill order, element ids, trajectories and find
    counters as the oracle's."""
    import dataclasses
    from pyrticle_amd import device as dv
    wl = dataclasses.replace(square, dt=square.dt * dt_scale)
    dm = dv.DeviceMesh(wl.ma)
    el = np.array([O.find_containing_element(oracle_square, xi) for xi in wl.x],
                  dtype=np.uint32)
    steps, n0 = (30 if dt_scale == 1.0 else 100), wl.x.shape[0]
    ox, op, oq, om_, oel = wl.x.copy(), wl.p.copy(), wl.q.copy(), wl.m.copy(), el.copy()
    oids = np.arange(n0, dtype=np.int64)
    n_o, cnt_o = O.lsrk4_run(oracle_square, wl.fields, wl.c, ox, op, oq, om_, oel, wl.dt,
                             steps, ids=oids)
    dm.set_static_fields([None if f is None else _dev(f, cuda) for f in wl.fields])
    tx, tp, tq, tm = (_dev(a, cuda) for a in (wl.x, wl.p, wl.q, wl.m))
    tel = dv.element_tensor(el, cuda)
    tids = torch.arange(n0, dtype=torch.int64, device=cuda)
    counters = torch.zeros(4, dtype=torch.int64, device=cuda)
    nrm = O.shape_normalizer(wl.radius, 2, wl.alpha)
    n = n0
    for _ in range(steps):
        dm.deposit(wl.radius, wl.alpha, nrm, tx[:n], tq[:n], tel[:n])
        n = dm.lsrk4_steps(1, wl.dt, wl.c, n, tx, tp, tq, tm, tel, ids=tids, counters=counters)
    assert n == n_o
    if dt_scale > 1:
        assert n < n0
    assert np.array_equal(tids[:n].cpu().numpy(), oids[:n_o])
    assert np.array_equal(dv.element_numpy(tel[:n]), oel[:n_o])
    assert rel_l2(tx[:n].cpu().numpy(), ox[:n_o]) < FP64_TOL
    assert rel_l2(tp[:n].cpu().numpy(), op[:n_o]) < FP64_TOL
    assert tuple(int(c) for c in counters.cpu()) == cnt_o


@pytest.fixture(scope="module")
def dense_p4():
    """Small P4 cylinder (1,368 tets, Np=35) filled with ~100 particles per
    element, so that most waves of the element-sorted beam sit in one element."""
    from pyrticle_amd import workloads as W
    mesh = W.cylinder_mesh_3d(n_rings=3, n_layers=8, order=4)
    return W.cylinder_dense_p4(n_particles=150_000, n_rings=3, mesh=mesh)


def test_lsrk4_fp32c_mfma_interpolation(dense_p4, cuda):
    """FP32C push of a dense P4 beam with non-uniform fields: the matrix-core
    interpolation (full waves inside one element) against the FP64 oracle and
    against the FP32 VALU interpolation; within the FP32 budget, bitwise run-to-
    run stable, and the matrix-core path carries most particle-stages."""
    from pyrticle_amd import device as dv
    wl = dense_p4
    om = O.OracleMesh(wl.ma)
    dm = dv.DeviceMesh(wl.ma)
    el = dv.element_numpy(dm.find_containing_element(_dev(wl.x, cuda)))
    assert np.all(el != dv.INVALID_ELEMENT)
    order = np.argsort(el, kind="stable")
    x0, p0, q0, m0 = (np.ascontiguousarray(a[order]) for a in (wl.x, wl.p, wl.q, wl.m))
    el0 = el[order]
    assert np.bincount(el0).mean() > 64
    steps = 2
    # non-uniform static fields, polynomials of degree <= 4 (exact in P4), so
    # that every monomial of the contraction contributes
    X = np.asarray(wl.ma.nodes).reshape(-1, 3)
    xx, yy, zz = X[:, 0], X[:, 1], X[:, 2]
    fields = [1e5 * (xx + 0.5 * yy * zz), 1e5 * (yy - 0.3 * xx ** 2),
              1e6 * (1 + 0.1 * zz - 0.02 * zz ** 2 + 0.05 * xx * yy),
              0.05 * xx * zz, -0.05 * yy * zz, 0.5 * (1 + 0.1 * xx ** 2 - 0.05 * yy ** 3)]
    ox, op, oq, om_, oel = x0.copy(), p0.copy(), q0.copy(), m0.copy(), el0.copy()
    n_o, _ = O.lsrk4_run(om, fields, wl.c, ox, op, oq, om_, oel, wl.dt, steps)
    dm.set_static_fields([_dev(f, cuda) for f in fields])
    out = {}
    try:
        for name, on in (("mfma", True), ("valu", False), ("mfma2", True)):
            dm.set_push_mfma(on, count=True)
            dm.push_mfma_waves()
            tx, tp, tq, tm = (_dev(a, cuda) for a in (x0, p0, q0, m0))
            tel = dv.element_tensor(el0, cuda)
            n = dm.lsrk4_steps(steps, wl.dt, wl.c, x0.shape[0], tx, tp, tq, tm, tel,
                               precision=dv.FP32C)
            assert n == n_o
            assert np.array_equal(dv.element_numpy(tel[:n]), oel[:n_o])
            out[name] = (tx[:n].cpu().numpy(), tp[:n].cpu().numpy(), dm.push_mfma_waves())
    finally:
        dm.set_push_mfma(False)
    # displacement and momentum change over the run against the FP64 oracle:
    # FP32 fields (relative rounding ~1e-7) enter through the momentum change
    dx_ref, dp_ref = ox[:n_o] - x0[:n_o], op[:n_o] - p0[:n_o]
    for name in ("mfma", "valu"):
        xs, ps, _ = out[name]
        assert rel_l2(xs - x0[:n_o], dx_ref) < 1e-6, name
        assert rel_l2(ps - p0[:n_o], dp_ref) < 1e-5, name
    assert np.array_equal(out["mfma"][1], out["mfma2"][1])      # deterministic
    # full waves (all but the last, and waves with a particle that left the
    # mesh at an earlier stage) take the matrix-core path, and each rank-1 step
    # rounds like the FMA chain of the VALU contraction: identical bits
    assert out["mfma"][2] >= 0.8 * 5 * steps * n_o and out["valu"][2] == 0
    assert np.array_equal(out["mfma"][1], out["valu"][1])
