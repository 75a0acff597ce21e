"""JStokes4th (src/current.cu:413-735) in the oracle, pinned by the
mathematics: under an affine map every octilateral is planar, its area
vector (corner triangles plus the inner diamond) is exact and the
trapezoid circulation of a linear field is exact, so J equals the curl of
a linear B to rounding; a uniform B gives 0; a smooth field on a smooth
map converges to its curl as h falls."""
import numpy as np

import oracle as o
from test_oracle_kat import grid, inn3, lattice, ref


def run(p, x, B):
    J = np.zeros(3 * p.ninner)
    o.lib.or_jstokes4th(ref(p), o.dp(np.ascontiguousarray(x).reshape(-1)), o.dp(np.ascontiguousarray(B).reshape(-1)),
                        o.dp(J))
    return inn3(p, J)


def curl_of(G):
    return np.array([G[2, 1] - G[1, 2], G[0, 2] - G[2, 0], G[1, 0] - G[0, 1]])


def test_linear_field_on_affine_grid_is_exact():
    p = grid(nx=6, ny=5, nz=7)
    rng = np.random.default_rng(31)
    A = np.eye(3) + 0.25 * rng.standard_normal((3, 3))
    G = rng.standard_normal((3, 3))
    x = lattice(p) @ A.T + 0.2
    B = x @ G.T + rng.standard_normal(3)
    J = run(p, x, B)
    want = curl_of(G)
    np.testing.assert_allclose(J, np.broadcast_to(want, J.shape), rtol=1e-10, atol=1e-11 * np.abs(want).max())


def test_uniform_field_gives_zero():
    p = grid(nx=5, ny=6, nz=4)
    rng = np.random.default_rng(32)
    x = lattice(p)
    x[1:-1, 1:-1, 1:-1] += 0.05 * rng.standard_normal(x[1:-1, 1:-1, 1:-1].shape)
    B = np.broadcast_to(rng.standard_normal(3), x.shape).copy()
    np.testing.assert_allclose(run(p, x, B), 0, atol=1e-12)


def test_smooth_field_converges_to_its_curl():
    alpha, kw = 0.06, 1.1
    errs = []
    for n in (9, 17):
        p = grid(nx=n, ny=n, nz=n, L=(2.0, 2.0, 2.0), O=(-1.0, -1.0, -1.0))
        X = lattice(p)
        x = X + alpha * np.stack([np.sin(kw * X[..., 1]) * np.sin(kw * X[..., 2]),
                                  np.sin(kw * X[..., 2]) * np.cos(kw * X[..., 0]),
                                  np.sin(kw * X[..., 0]) * np.sin(kw * X[..., 1])], -1)
        u, v, w = x[..., 0], x[..., 1], x[..., 2]
        B = np.stack([np.sin(v) + 0.3 * w, np.cos(w) * 0.5, np.sin(u) * np.cos(v)], -1)
        # curl of B in x: (dBz/dy - dBy/dz, dBx/dz - dBz/dx, dBy/dx - dBx/dy)
        want = np.stack([-np.sin(u) * np.sin(v) + 0.5 * np.sin(w), 0.3 - np.cos(u) * np.cos(v), -np.cos(v)], -1)
        J = run(p, x, B)
        errs.append(float(np.abs(J - want[1:-1, 1:-1, 1:-1]).max()))
    assert errs[1] < errs[0] / 3.2, errs
    print("JStokes4th errors", errs)
