"""Parameter sets of the braid relaxation workloads.

braid_params(n) builds the divergence-free braid of six twisted blobs
    B0 = z^ + sum_k c_k (-(y-y_k), x-x_k, 0) exp(-((x-x_k)^2+(y-y_k)^2)/2 - (z-z_k)^2/4)
on the box [-4,4] x [-4,4] x [-24,24], c_k = +1,-1,..., x_k = +1,-1,...,
y_k = 0, z_k = -20,-12,-4,4,12,20, line-tied on all faces, with the
reference's "blobs" field (src/init.cu:109-124): c_k = 2 ampl twist/scale,
scale^2 = 2, zl^2 = 4.
"""
from __future__ import annotations

import numpy as np

from . import Params, params_default, params_derive, braid_seeded, init_state

SCALE = 1.4142135623730951       # sqrt(2): exponent (...)/2
TWIST = 0.7071067811865476       # c * scale / 2 with |c| = 1
GRIDS = {
    "small": (33, 33, 97),
    "medium": (129, 129, 385),
    "large": (257, 257, 769),
}


def _box(p, nx, ny, nz):
    p.nx, p.ny, p.nz = nx, ny, nz
    p.Lx, p.Ly, p.Lz = 8.0, 8.0, 48.0
    p.Ox, p.Oy, p.Oz = -4.0, -4.0, -24.0


def braid_params(nx=33, ny=33, nz=97, n_cycle=200, max_error=1e-5):
    """Plain braid, friction 1 (epsilonProf off), pressure off, adaptive
    Cash-Karp with tolerance max_error; dt0/dtMin keep the reference
    defaults (src/readNamelist.f90:115-116)."""
    p = params_default()
    _box(p, nx, ny, nz)
    p.set_str("bInit", "blobs")
    p.ampl = 1.0
    p.nBlobs = 6
    for k in range(6):
        s = 1.0 if k % 2 == 0 else -1.0
        p.blobXc[k] = s
        p.blobYc[k] = 0.0
        p.blobZc[k] = -20.0 + 8.0 * k
        p.blobScale[k] = SCALE
        p.blobZl[k] = 2.0
        p.blobTwist[k] = s * TWIST
    p.nCycle = n_cycle
    p.maxError = max_error
    p.pressure = False
    p.epsilonProf = False
    p.zUpdate = True
    p.set_str("jMethod", "Stokes")
    p.nTs = 1
    p.redB2 = p.redJxB_B2Max = p.redJMax = True
    return params_derive(p)


def small():
    """33 x 33 x 97 braid, 200 adaptive steps at tolerance 1e-5."""
    return braid_params(*GRIDS["small"], n_cycle=200)


def medium():
    """129 x 129 x 385 braid, run until max|JxB|/(|J||B|) <= 1e-3."""
    return braid_params(*GRIDS["medium"], n_cycle=10 ** 8)


def large(seed=3):
    """257 x 257 x 769 braid with seeded amplitudes c_k = (-1)^k U(0.5,1.5),
    pressure on with beta = 0.1 mean(|B0|^2) over inner points, 2000 steps."""
    p = braid_params(*GRIDS["large"], n_cycle=2000)
    braid_seeded(p, seed, 6, 6, False)
    p.pressure = True
    p.set_str("pMethod", "Classic")
    p.beta = 0.1 * mean_b0_sq(p)
    return p


def mean_b0_sq(p):
    """Mean of |B0|^2 over inner points of the blobs field, evaluated on the
    host from the closed form."""
    B0, _ = init_state(p)
    b = B0.reshape(p.nz + 2, p.ny + 2, p.nx + 2, 3)
    s = 0.0
    for k in range(1, p.nz + 1):          # plane by plane: no full-size temporaries
        q = b[k, 1:-1, 1:-1]
        s += float(np.einsum("ijl,ijl->", q, q))
    return s / (p.nx * p.ny * p.nz)


def ensemble_member(i, sizes=((33, 33, 97), (49, 49, 145), (65, 65, 193), (97, 97, 289))):
    """Member i of the seeded ensemble: grid size and 2-8 blobs from seed 1000+i."""
    rng = np.random.default_rng(1000 + i)
    nx, ny, nz = sizes[int(rng.integers(len(sizes)))]
    p = braid_params(nx, ny, nz, n_cycle=10 ** 8)
    braid_seeded(p, 1000 + i, 2, 8, True)
    return p
