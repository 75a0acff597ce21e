import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    config.addinivalue_line("markers", "slow: longer CPU test")


def gpu_available():
    """A ROCm device node is present (checked without initialising HIP)."""
    return os.path.exists("/dev/kfd")


@pytest.fixture(scope="session")
def require_gpu():
    if not gpu_available():
        pytest.skip("no HIP device")


def small_braid(nx=17, ny=19, nz=23, **kw):
    """A small braid on a non-cubic grid (distinct extents catch transposed
    indices), blobs scaled into the box."""
    from glemur_amd import workloads, params_derive
    p = workloads.braid_params(nx, ny, nz, n_cycle=50)
    p.Lx, p.Ly, p.Lz = 4.0, 4.5, 12.0
    p.Ox, p.Oy, p.Oz = -2.0, -2.25, -6.0
    for k in range(6):
        p.blobXc[k] = 0.5 * (1 if k % 2 == 0 else -1)
        p.blobZc[k] = -5.0 + 2.0 * k
        p.blobScale[k] = 0.8
        p.blobZl[k] = 1.0
        p.blobTwist[k] = (1 if k % 2 == 0 else -1) * 0.4
    for key, v in kw.items():
        setattr(p, key, v)
    return params_derive(p)


def evolved_state(p, nsteps):
    """(B0, xb) after nsteps oracle steps from the initial state, so kernels
    see a distorted grid."""
    import oracle as o
    q = o.from_product(p)
    B0, xb = o.init_state(q)
    if nsteps == 0:
        return B0, xb
    sim = o.Sim(q, B0, xb)
    for _ in range(nsteps):
        sim.step()
    return B0.copy(), sim.arr("xb").copy()


def inner_mask(p):
    m = np.zeros((p.nz + 2, p.ny + 2, p.nx + 2), dtype=bool)
    m[1:-1, 1:-1, 1:-1] = True
    return m.reshape(-1)


def edge_mask(p):
    """Ghost points with two or three ghost coordinates (edges, corners)."""
    k, j, i = np.meshgrid(np.arange(p.nz + 2), np.arange(p.ny + 2), np.arange(p.nx + 2), indexing="ij")
    g = ((i == 0) | (i == p.nx + 1)).astype(int) + ((j == 0) | (j == p.ny + 1)).astype(int) + \
        ((k == 0) | (k == p.nz + 1)).astype(int)
    return (g >= 2).reshape(-1)


def face_mask(p):
    """Ghost points with exactly one ghost coordinate (the six faces)."""
    k, j, i = np.meshgrid(np.arange(p.nz + 2), np.arange(p.ny + 2), np.arange(p.nx + 2), indexing="ij")
    g = ((i == 0) | (i == p.nx + 1)).astype(int) + ((j == 0) | (j == p.ny + 1)).astype(int) + \
        ((k == 0) | (k == p.nz + 1)).astype(int)
    return (g == 1).reshape(-1)
