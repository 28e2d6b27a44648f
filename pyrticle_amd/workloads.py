"""Seeded synthetic workloads of the particle path.

* ``square_2d``: [-1,1]^2 from a 33x33 jittered grid (2,048 triangles), P3
  (Np=10); 10,000 electrons uniform in a disk of radius 0.5 with momenta
  Gaussian around gamma=2 along x (1% relative spread); R=0.1, alpha=2;
  static E=(0,1e3,0) V/m, B=(0,0,0.01) T (TE fields ex, ey, bz);
  dt = 0.2*h_min/c; 100 steps.
* ``cylinder_3d``: cylinder of radius 1, z in [0,4], ~200k P3 tetrahedra
  (Np=20); 4M electrons, transverse uniform in an ellipse with semi-axes
  0.3 x 0.2, z ~ N(2, 0.4^2) (re-drawn inside the mesh), gamma=10 along z with
  0.1% momentum spread; R=0.08, alpha=2; static solenoid B_z=0.5 T plus a
  uniform E_z=1e6 V/m; dt = 0.2*h_min/c; 200 steps.
* ``cylinder_dense_p4``: the dense-tile variant: the same cylinder with 49,896
  P4 tetrahedra (Np=35, 7 rings x 54 layers) and 16M electrons uniform in the
  cylinder inscribed in the mesh (about 320 per element), same momenta, fields,
  R and dt rule.

Random streams use numpy.random.default_rng: seed 0 for meshes, seed 1 for
particles.  Particles that land outside the mesh are re-drawn (the reference's
add_particles would silently drop them, cloud.py:470-474).
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import discretization as D
from .meshdata import MeshArrays, build_mesh_arrays
from .units import SIUnitsWithNaturalConstants as U


@dataclass
class Workload:
    name: str
    discr: D.Discretization
    ma: MeshArrays
    x: np.ndarray
    p: np.ndarray
    q: np.ndarray
    m: np.ndarray
    radius: float
    alpha: float
    fields: list          # six nodal arrays or None (ex, ey, ez, bx, by, bz)
    dt: float
    nsteps: int
    c: float


def _momenta_gamma(gamma, n, axis, spread, rng, dim):
    c = U.VACUUM_LIGHT_SPEED()
    beta = (1 - 1 / gamma ** 2) ** 0.5
    p0 = gamma * U.EL_MASS * beta * c
    p = rng.normal(0.0, spread * p0, size=(n, dim))
    p[:, axis] += p0
    return p


def square_2d(n_particles: int = 10_000, order: int = 3, seed: int = 1,
              grid: int = 33) -> Workload:
    mesh = D.make_jittered_square_mesh(n=grid, seed=0)
    discr = D.Discretization(mesh, order)
    ma = build_mesh_arrays(discr)
    rng = np.random.default_rng(seed)
    r = 0.5 * np.sqrt(rng.uniform(size=n_particles))
    phi = rng.uniform(0, 2 * np.pi, size=n_particles)
    x = np.stack([r * np.cos(phi), r * np.sin(phi)], axis=1)
    p = _momenta_gamma(2.0, n_particles, 0, 0.01, rng, 2)
    q = np.full(n_particles, -U.EL_CHARGE)
    m = np.full(n_particles, U.EL_MASS)
    nn = len(discr)
    fields = [np.zeros(nn), np.full(nn, 1e3), None, None, None, np.full(nn, 0.01)]
    c = U.VACUUM_LIGHT_SPEED()
    dt = 0.2 * D.min_edge_length(mesh) / c
    return Workload("square_2d", discr, ma, x, p, q, m, 0.1, 2.0, fields, dt, 100, c)


def cylinder_mesh_3d(n_rings: int = 12, n_layers: int = 74, order: int = 3):
    mesh = D.make_cylinder_mesh(radius=1.0, z0=0.0, z1=4.0, n_rings=n_rings,
                                n_layers=n_layers, seed=0)
    discr = D.Discretization(mesh, order)
    return discr, build_mesh_arrays(discr)


def cylinder_beam(n_particles: int, seed: int = 1, axes=(0.3, 0.2), sigma_z: float = 0.4,
                  gamma: float = 10.0, spread: float = 1e-3):
    rng = np.random.default_rng(seed)
    r = np.sqrt(rng.uniform(size=n_particles))
    phi = rng.uniform(0, 2 * np.pi, size=n_particles)
    z = rng.normal(2.0, sigma_z, size=n_particles)
    bad = (z <= 0.0) | (z >= 4.0)
    while bad.any():
        z[bad] = rng.normal(2.0, sigma_z, size=int(bad.sum()))
        bad = (z <= 0.0) | (z >= 4.0)
    x = np.stack([axes[0] * r * np.cos(phi), axes[1] * r * np.sin(phi), z], axis=1)
    p = _momenta_gamma(gamma, n_particles, 2, spread, rng, 3)
    q = np.full(n_particles, -U.EL_CHARGE)
    m = np.full(n_particles, U.EL_MASS)
    return x, p, q, m


def cylinder_3d(n_particles: int = 4_000_000, n_rings: int = 12, n_layers: int = 74,
                order: int = 3, seed: int = 1, mesh=None) -> Workload:
    discr, ma = mesh if mesh is not None else cylinder_mesh_3d(n_rings, n_layers, order)
    x, p, q, m = cylinder_beam(n_particles, seed)
    nn = len(discr)
    fields = [np.zeros(nn), np.zeros(nn), np.full(nn, 1e6),
              np.zeros(nn), np.zeros(nn), np.full(nn, 0.5)]
    c = U.VACUUM_LIGHT_SPEED()
    dt = 0.2 * D.min_edge_length(discr.mesh) / c
    return Workload("cylinder_3d", discr, ma, x, p, q, m, 0.08, 2.0, fields, dt, 200, c)


def cylinder_dense_p4(n_particles: int = 16_000_000, n_rings: int = 7, n_layers: int = 54,
                      order: int = 4, seed: int = 1, mesh=None) -> Workload:
    discr, ma = mesh if mesh is not None else cylinder_mesh_3d(n_rings, n_layers, order)
    rng = np.random.default_rng(seed)
    # uniform in the cylinder inscribed in the mesh's polygonal cross-section
    # (the outer ring has round(2 pi n_rings) vertices on the unit circle)
    r_in = 0.99 * np.cos(np.pi / max(6, int(round(2 * np.pi * n_rings))))
    r = r_in * np.sqrt(rng.uniform(size=n_particles))
    phi = rng.uniform(0, 2 * np.pi, size=n_particles)
    z = rng.uniform(1e-3, 4.0 - 1e-3, size=n_particles)
    x = np.stack([r * np.cos(phi), r * np.sin(phi), z], axis=1)
    p = _momenta_gamma(10.0, n_particles, 2, 1e-3, rng, 3)
    q = np.full(n_particles, -U.EL_CHARGE)
    m = np.full(n_particles, U.EL_MASS)
    nn = len(discr)
    fields = [np.zeros(nn), np.zeros(nn), np.full(nn, 1e6),
              np.zeros(nn), np.zeros(nn), np.full(nn, 0.5)]
    c = U.VACUUM_LIGHT_SPEED()
    dt = 0.2 * D.min_edge_length(discr.mesh) / c
    return Workload("cylinder_dense_p4", discr, ma, x, p, q, m, 0.08, 2.0, fields, dt, 200, c)
