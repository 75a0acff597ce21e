"""JClassic (src/current.cu:10-219) in the oracle, pinned by the
mathematics: J is curl_x of B = jac B0 / det.  Under an affine map with a
linear B0 every central difference is exact, so J equals the closed-form
curl of (A M A^-1) / det A; under a smooth nonlinear map with a smooth B0
the error against automatic differentiation (torch autograd, an independent
route) falls like h^2."""
import numpy as np
import pytest

import oracle as o
from test_oracle_kat import grid, inn3, lattice, ref


def run_jclassic(p, xb, B0, det):
    J = np.zeros(3 * p.ninner)
    o.lib.or_jclassic(ref(p), o.dp(np.ascontiguousarray(xb).reshape(-1)), o.dp(np.ascontiguousarray(B0).reshape(-1)),
                      o.dp(np.ascontiguousarray(det).reshape(-1)), o.dp(J))
    return inn3(p, J)


def test_affine_map_linear_field_closed_form():
    p = grid(nx=6, ny=5, nz=7)
    rng = np.random.default_rng(21)
    A = np.eye(3) + 0.25 * rng.standard_normal((3, 3))
    M = rng.standard_normal((3, 3))
    b0 = rng.standard_normal(3)
    X = lattice(p)
    x = X @ A.T + 0.3
    B0 = X @ M.T + b0
    dA = np.linalg.det(A)
    det = np.full(X.shape[:3], dA)
    G = A @ M @ np.linalg.inv(A) / dA          # dB_k / dx_j
    want = np.array([G[2, 1] - G[1, 2], G[0, 2] - G[2, 0], G[1, 0] - G[0, 1]])
    J = run_jclassic(p, x, B0, det)
    np.testing.assert_allclose(J, np.broadcast_to(want, J.shape), rtol=1e-10, atol=1e-11 * np.abs(want).max())


def _smooth(Xt, alpha, k):
    import torch
    X, Y, Z = Xt[..., 0], Xt[..., 1], Xt[..., 2]
    x = torch.stack([X + alpha * torch.sin(k * Y) * torch.sin(k * Z),
                     Y + alpha * torch.sin(k * Z) * torch.cos(k * X),
                     Z + alpha * torch.sin(k * X) * torch.sin(k * Y)], -1)
    B0 = torch.stack([0.4 * torch.sin(Y), 0.3 * torch.cos(Z) * torch.sin(X), 1 + 0.2 * torch.sin(X + Y)], -1)
    return x, B0


def test_smooth_map_converges_to_the_curl():
    import torch
    from torch.func import jacrev, vmap

    alpha, kw = 0.07, 1.3

    def bfield(X):
        jac = jacrev(lambda v: _smooth(v, alpha, kw)[0])(X)
        return jac @ _smooth(X, alpha, kw)[1] / torch.linalg.det(jac)

    errs = []
    for n in (9, 17):
        p = grid(nx=n, ny=n, nz=n, L=(2.0, 2.0, 2.0), O=(-1.0, -1.0, -1.0))
        X = torch.tensor(lattice(p))
        with torch.no_grad():
            x, B0 = _smooth(X, alpha, kw)
            det = vmap(lambda v: torch.linalg.det(jacrev(lambda w: _smooth(w, alpha, kw)[0])(v)))(X.reshape(-1, 3))
        J = run_jclassic(p, x.numpy(), B0.numpy(), det.numpy())
        Xi = X[1:-1, 1:-1, 1:-1].reshape(-1, 3)
        dBdX = vmap(jacrev(bfield))(Xi)
        jac = vmap(jacrev(lambda v: _smooth(v, alpha, kw)[0]))(Xi)
        G = dBdX @ torch.linalg.inv(jac)           # dB_k / dx_j
        want = torch.stack([G[:, 2, 1] - G[:, 1, 2], G[:, 0, 2] - G[:, 2, 0], G[:, 1, 0] - G[:, 0, 1]], -1)
        errs.append(float(np.abs(J.reshape(-1, 3) - want.detach().numpy()).max()))
        assert float(want.abs().max()) > 0.1
    assert errs[1] < errs[0] / 3.2, errs
    assert errs[1] < 5e-3, errs
