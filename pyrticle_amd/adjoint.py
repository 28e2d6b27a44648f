"""Differentiable push + deposition: the adjoint kernels under torch.autograd.

pyrticle has no adjoint; this module differentiates its particle path -- the
fused LSRK4 push in static fields (cloud.py:554-603 with push_monomial.hpp
forces) followed by shape-function deposition (dep_shape.hpp:67-150) -- with
hand-written backward kernels:

* forward: ``pyr_lsrk4_forward`` runs all steps without compaction (a particle
  that leaves the mesh keeps its slot and is marked dead) and stores the
  particle state every ``ckpt_every`` steps; then ``pyr_deposit`` of the final
  state;
* backward: ``pyr_deposit_backward`` (per-particle gather of grad S times the
  adjoint densities), ``pyr_velocities_backward`` for the current's velocity
  dependence, and ``pyr_lsrk4_adjoint``, which recomputes each checkpoint
  segment and runs the five LSRK4 stages of every step in reverse.

Particles that leave the mesh get zero gradients.  Element assignment is
piecewise constant and has no derivative; fields enter through their DG
interpolant, which is differentiated inside each element.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np
import torch

from . import device as dv
from .units import SIUnitsWithNaturalConstants as U


@dataclass
class PushDepositConfig:
    """Everything the differentiable map (x0, p0) -> (rho, j) holds fixed."""
    dmesh: dv.DeviceMesh
    fields: list                 # six nodal device arrays or None (ex, ey, ez, bx, by, bz)
    q: torch.Tensor              # n, float64
    m: torch.Tensor              # n, float64
    elem0: torch.Tensor          # n, int32 (uint32 bits): containing elements of x0
    nsteps: int
    dt: float
    c: float
    radius: float
    alpha: float
    normalizer: float
    precision: int = dv.FP64
    ckpt_every: int = 20
    want_j: bool = False
    stats: dict = field(default_factory=dict)


class PushDepositFunction(torch.autograd.Function):
    """(x0, p0) -> (rho, j) after ``nsteps`` LSRK4 steps; j is empty unless
    ``cfg.want_j``."""

    @staticmethod
    def forward(ctx, x0, p0, cfg: PushDepositConfig):
        dm = cfg.dmesh
        n, d = x0.shape
        x = x0.detach().contiguous().clone()
        p = p0.detach().contiguous().clone()
        elem = cfg.elem0.clone()
        death = torch.zeros(n, dtype=torch.int32, device=x.device)
        nseg = max(1, math.ceil(cfg.nsteps / cfg.ckpt_every))
        ckpt_xp = torch.empty((nseg, n, 2 * d), dtype=torch.float64, device=x.device)
        ckpt_e = torch.empty((nseg, n), dtype=torch.int32, device=x.device)
        dm.set_static_fields(cfg.fields)
        dm.lsrk4_forward(cfg.nsteps, cfg.dt, cfg.c, x, p, cfg.q, cfg.m, elem, death,
                         cfg.ckpt_every, ckpt_xp, ckpt_e, precision=cfg.precision)
        v = dv.velocities(p, cfg.m, cfg.c)
        rho, j = dm.deposit(cfg.radius, cfg.alpha, cfg.normalizer, x, cfg.q, elem, v=v,
                            want_rho=True, want_j=cfg.want_j, precision=cfg.precision)
        ctx.cfg = cfg
        ctx.save_for_backward(x, p, v, elem, death, ckpt_xp, ckpt_e)
        cfg.stats["alive"] = int((death == 0).sum().item())
        if j is None:
            j = torch.zeros((0, d), dtype=torch.float64, device=x.device)
        if not cfg.want_j:
            ctx.mark_non_differentiable(j)
        return rho, j

    @staticmethod
    def backward(ctx, rho_bar, j_bar):
        cfg: PushDepositConfig = ctx.cfg
        dm = cfg.dmesh
        x, p, v, elem, death, ckpt_xp, ckpt_e = ctx.saved_tensors
        rb = rho_bar.contiguous() if rho_bar is not None else None
        jb = j_bar.contiguous() if (cfg.want_j and j_bar is not None) else None
        x_bar, v_bar = dm.deposit_backward(cfg.radius, cfg.alpha, cfg.normalizer, x, cfg.q,
                                           elem, rho_bar=rb, j_bar=jb, v=v,
                                           precision=cfg.precision)
        if v_bar is not None:
            p_bar = dv.velocities_backward(p, cfg.m, cfg.c, v_bar)
        else:
            p_bar = torch.zeros_like(p)
        dm.set_static_fields(cfg.fields)
        dm.lsrk4_adjoint(cfg.nsteps, cfg.ckpt_every, cfg.dt, cfg.c, cfg.q, cfg.m, death,
                         ckpt_xp, ckpt_e, x_bar, p_bar, precision=cfg.precision)
        return x_bar, p_bar, None


def push_and_deposit(x0, p0, cfg: PushDepositConfig):
    """Differentiable (rho, j) of the particles (x0, p0) after cfg.nsteps steps."""
    return PushDepositFunction.apply(x0, p0, cfg)


# ---------------------------------------------------------------------------
# differentiable beam
# ---------------------------------------------------------------------------
BEAM_PARAMETERS = ("semi_axis_x", "semi_axis_y", "sigma_z", "gamma", "spread_transverse",
                   "spread_longitudinal")


@dataclass
class BeamNoise:
    """Fixed random draws of a KV-like beam; the beam parameters act on them
    smoothly, so the map parameters -> (x0, p0) is differentiable."""
    u: torch.Tensor        # radius^2 in the unit disk
    phi: torch.Tensor
    xi_z: torch.Tensor     # standard normal, |xi_z| < z_cut
    xi_p: torch.Tensor     # n x 3 standard normal
    z_center: float


def beam_noise(n: int, seed: int, z_center: float = 2.0, z_cut: float = 4.5,
               device="cuda") -> BeamNoise:
    rng = np.random.default_rng(seed)
    u = rng.uniform(size=n)
    phi = rng.uniform(0.0, 2 * np.pi, size=n)
    xi_z = rng.normal(size=n)
    bad = np.abs(xi_z) >= z_cut
    while bad.any():
        xi_z[bad] = rng.normal(size=int(bad.sum()))
        bad = np.abs(xi_z) >= z_cut
    xi_p = rng.normal(size=(n, 3))
    t = lambda a: torch.as_tensor(a, dtype=torch.float64, device=device)
    return BeamNoise(t(u), t(phi), t(xi_z), t(xi_p), z_center)


def beam_state(theta: torch.Tensor, noise: BeamNoise, mass: float = U.EL_MASS,
               c: float | None = None):
    """(x0, p0) of the beam with parameters theta = (a, b, sigma_z, gamma,
    spread_t, spread_l): transverse uniform in the ellipse with semi-axes a, b;
    z = z_center + sigma_z xi_z; p = m c sqrt(gamma^2 - 1) (s_t xi_x, s_t xi_y,
    1 + s_l xi_z')."""
    if c is None:
        c = U.VACUUM_LIGHT_SPEED()
    a, b, sz, gamma, st, sl = theta.unbind(0)
    r = torch.sqrt(noise.u)
    x = torch.stack([a * r * torch.cos(noise.phi), b * r * torch.sin(noise.phi),
                     noise.z_center + sz * noise.xi_z], dim=1)
    p0 = mass * c * torch.sqrt(gamma * gamma - 1.0)
    p = p0 * torch.stack([st * noise.xi_p[:, 0], st * noise.xi_p[:, 1],
                          1.0 + sl * noise.xi_p[:, 2]], dim=1)
    return x, p


def default_beam_theta(device="cuda"):
    """Parameters of the cylinder workload's beam."""
    return torch.tensor([0.3, 0.2, 0.4, 10.0, 1e-3, 1e-3], dtype=torch.float64, device=device)
