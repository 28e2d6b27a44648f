import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP) GPU")
    config.addinivalue_line("markers", "slow: long-running CPU test")


@pytest.fixture(scope="session")
def square():
    """The 2D workload (2,048 P3 triangles, 10k electrons)."""
    from pyrticle_amd.workloads import square_2d
    return square_2d()


@pytest.fixture(scope="session")
def oracle_square(square):
    from oracle import oracle as O
    return O.OracleMesh(square.ma)


@pytest.fixture(scope="session")
def small_cylinder():
    """Small 3D P3 cylinder (~5k tets) with a 20k-particle beam."""
    from pyrticle_amd import workloads as W
    discr, ma = W.cylinder_mesh_3d(n_rings=5, n_layers=12, order=3)
    return W.cylinder_3d(n_particles=20_000, mesh=(discr, ma))


@pytest.fixture(scope="session")
def periodic_cylinder():
    from pyrticle_amd import discretization as D
    from pyrticle_amd.meshdata import build_mesh_arrays
    mesh = D.make_cylinder_mesh(radius=1.0, z0=0.0, z1=2.0, n_rings=4, n_layers=8,
                                periodic=True, seed=0)
    discr = D.Discretization(mesh, 1)
    return discr, build_mesh_arrays(discr)


@pytest.fixture(scope="session")
def cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test collected but no GPU is visible")
    return torch.device("cuda:0")


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))
