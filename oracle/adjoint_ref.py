"""FP64 CPU reference for gradients of the particle path (torch autograd).

TEST INFRASTRUCTURE ONLY: tests/ import this module as the checker of the
adjoint kernels.  The shipped package never imports it.

The reference has no adjoint.  This module restates its forward path in
plain torch float64 on the CPU, differentiable by autograd, independently of
the HIP kernels:

* push: hedge's LSRK4 (Carpenter & Kennedy 1994, the published 2N-storage
  coefficients) of the particle RHS [v(p), q E(x) + v x q B(x)]
  (push_monomial.hpp:265-305, particle_state.hpp:98-125), with the fields
  given as a function of position -- for polynomial fields of degree <= N
  this equals the DG interpolant the reference evaluates (exact in every
  element); localization and kills are not modelled, so callers check that
  no particle leaves the mesh;
* deposition: S = N (R - r^2/R)^alpha over every (particle, node) pair within
  R (tools.hpp:514-534, dep_shape.hpp:88-106), pairs from a k-d tree over the
  DG nodes, rho_i = sum q S and j_ia = sum q v_a S (dep_target.hpp:54-186).
"""

from __future__ import annotations

import numpy as np
import torch
from scipy.spatial import cKDTree

# Carpenter & Kennedy (1994), NASA TM-109112, five-stage 2N-storage RK4
RK4A = (0.0, -567301805773.0 / 1357537059087.0, -2404267990393.0 / 2016746695238.0,
        -3550918686646.0 / 2091501179385.0, -1275806237668.0 / 842570457699.0)
RK4B = (1432997174477.0 / 9575080441755.0, 5161836677717.0 / 13612068292357.0,
        1720146321549.0 / 2090206949498.0, 3134564353537.0 / 4481467310338.0,
        2277821191437.0 / 14882151754819.0)


def velocities(p, m, c):
    """v = c p / sqrt(m^2 c^2 + |p|^2)."""
    pp = (p * p).sum(dim=1, keepdim=True)
    return c * p / torch.sqrt((m * m * c * c)[:, None] + pp)


def _cross(a, b):
    return torch.stack([a[:, 1] * b[:, 2] - a[:, 2] * b[:, 1],
                        a[:, 2] * b[:, 0] - a[:, 0] * b[:, 2],
                        a[:, 0] * b[:, 1] - a[:, 1] * b[:, 0]], dim=1)


def _pad3(a):
    if a.shape[1] == 3:
        return a
    return torch.cat([a, torch.zeros((a.shape[0], 3 - a.shape[1]), dtype=a.dtype)], dim=1)


def push(x, p, q, m, c, dt, nsteps, field_fn):
    """nsteps LSRK4 steps; field_fn(x) -> (E, B), each n x 3 (torch)."""
    d = x.shape[1]
    for _ in range(nsteps):
        rx = torch.zeros_like(x)
        rp = torch.zeros_like(p)
        for s in range(5):
            v = velocities(p, m, c)
            E, B = field_fn(x)
            F = q[:, None] * E + _cross(_pad3(v), q[:, None] * B)
            rx = RK4A[s] * rx + dt * v
            rp = RK4A[s] * rp + dt * F[:, :d]
            x = x + RK4B[s] * rx
            p = p + RK4B[s] * rp
    return x, p


class NodePairs:
    """(particle, node) pairs within R, from a k-d tree over the DG nodes."""

    def __init__(self, nodes: np.ndarray, radius: float):
        self.nodes = np.ascontiguousarray(nodes, dtype=np.float64)
        self.radius = radius
        self.tree = cKDTree(self.nodes)

    def pairs(self, x: np.ndarray):
        lists = self.tree.query_ball_point(np.asarray(x), self.radius)
        pi = np.repeat(np.arange(len(lists)), [len(l) for l in lists])
        ni = np.concatenate([np.asarray(l, dtype=np.int64) for l in lists]) if len(lists) \
            else np.zeros(0, np.int64)
        return torch.from_numpy(pi.astype(np.int64)), torch.from_numpy(ni)


def deposit(pairs: NodePairs, x, q, v, alpha, normalizer, want_j=True):
    """(rho, j) at the DG nodes; j is nodes x d."""
    R = pairs.radius
    pi, ni = pairs.pairs(x.detach().numpy())
    nodes = torch.from_numpy(pairs.nodes)
    diff = nodes[ni] - x[pi]
    r2 = (diff * diff).sum(dim=1)
    t = torch.clamp(R - r2 / R, min=0.0)
    S = normalizer * t ** alpha
    nn = nodes.shape[0]
    w = q[pi] * S
    rho = torch.zeros(nn, dtype=torch.float64).index_add(0, ni, w)
    j = None
    if want_j:
        j = torch.zeros((nn, x.shape[1]), dtype=torch.float64).index_add(
            0, ni, w[:, None] * v[pi])
    return rho, j


def beam_state(theta, u, phi, xi_z, xi_p, z_center, mass, c):
    """The differentiable beam of pyrticle_amd.adjoint.beam_state, restated."""
    a, b, sz, gamma, st, sl = theta.unbind(0)
    r = torch.sqrt(u)
    x = torch.stack([a * r * torch.cos(phi), b * r * torch.sin(phi), z_center + sz * xi_z],
                    dim=1)
    p0 = mass * c * torch.sqrt(gamma * gamma - 1.0)
    p = p0 * torch.stack([st * xi_p[:, 0], st * xi_p[:, 1], 1.0 + sl * xi_p[:, 2]], dim=1)
    return x, p
