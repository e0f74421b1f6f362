import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


# ---------------------------------------------------------------- shared workloads
# 2D lid-driven cavity: 100x100 lattice on [0,1]^2, interior jitter +-0.25h,
# PHS r^3 + degree-3 polynomial, 30-node stencils (the 10k-node cavity case).
CAVITY2D = dict(dim=2, n=100, jitter=0.25, seed=20240601, k_interior=30)
# Small 3D cube with the 1M-node case's stencil rule: 40 interior nodes,
# 60/55/50 on lattice layers 0/1/2 (faces and the two layers under them).
CUBE3D_SMALL = dict(dim=3, n=16, jitter=0.3, seed=7, k_interior=40, k_layers=(60, 55, 50))


@pytest.fixture(scope="session")
def cavity2d():
    import memphys_amd as M
    return M.box_cloud(**CAVITY2D)


@pytest.fixture(scope="session")
def cube3d_small():
    import memphys_amd as M
    return M.box_cloud(**CUBE3D_SMALL)


def lid_bc(cloud, lid_axis):
    """Cavity boundary conditions: velocity Dirichlet on every wall (u = 1 on
    the lid face x[lid_axis] = 1), pressure Neumann everywhere."""
    nv, dim = cloud.nv, cloud.dim
    u_dir = np.ones(nv, dtype=np.uint8)
    p_dir = np.zeros(nv, dtype=np.uint8)
    u0 = np.zeros(nv)
    lid = (cloud.boundary == 1) & (np.abs(cloud.xyz[:, lid_axis] - 1.0) < 1e-12)
    u0[lid] = 1.0
    return u_dir, p_dir, u0


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (den if den > 0 else 1.0)
