"""jxbAver (src/rungeKutta.cu:172-177, 357-443) in the oracle's kk: the
slope uses (1 - w) J x B plus w/6 times J x B of each face neighbour that is
an inner point.  Pinned by closed forms: w = 0 is the plain slope; a uniform
J x B keeps its value inside and loses w/6 per missing neighbour on faces,
edges and corners; a single nonzero point spreads w/6 to exactly its six
face neighbours."""
import numpy as np

import oracle as o
from test_oracle_kat import grid, inn3, pad3, ref


def slope(p, J, Bf):
    _, b, _, _ = o.coefficients()
    xb = np.zeros(3 * p.npad)
    G = np.zeros(3 * p.ninner)
    kk = np.zeros(18 * p.ninner)
    xt = xb.copy()
    o.lib.or_kk(ref(p), o.dp(b), 0, 1.0, o.dp(xb), o.dp(Bf), o.dp(J), o.dp(G), o.dp(kk), o.dp(xt))
    return inn3(p, kk[:3 * p.ninner]).copy()


def uniform(p):
    Bf = np.zeros(3 * p.npad)
    pad3(p, Bf)[..., 1] = 1.0                 # B = y^
    J = np.zeros(3 * p.ninner)
    inn3(p, J)[..., 0] = 2.0                  # J = 2 x^  ->  J x B = 2 z^
    return J, Bf


def test_zero_weight_is_the_plain_slope():
    rng = np.random.default_rng(41)
    p0 = grid(nx=5, ny=4, nz=6)
    p1 = grid(nx=5, ny=4, nz=6, jxbAver=True, jxbAverWeight=0.0)
    J = rng.standard_normal(3 * p0.ninner)
    Bf = rng.standard_normal(3 * p0.npad)
    assert np.array_equal(slope(p0, J, Bf), slope(p1, J, Bf))


def test_uniform_field_loses_w_over_6_per_missing_neighbour():
    w = 0.3
    p = grid(nx=5, ny=4, nz=6, jxbAver=True, jxbAverWeight=w)
    J, Bf = uniform(p)
    k = slope(p, J, Bf)[..., 2]
    nz, ny, nx = k.shape
    iz, iy, ix = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    missing = sum(((a == 0) | (a == n - 1)).astype(int) for a, n in ((ix, nx), (iy, ny), (iz, nz)))
    np.testing.assert_allclose(k, 2.0 * (1 - w * missing / 6), rtol=1e-15, atol=1e-15)


def test_a_point_spreads_to_its_six_face_neighbours():
    w = 0.42
    p = grid(nx=5, ny=6, nz=7, jxbAver=True, jxbAverWeight=w)
    Bf = np.zeros(3 * p.npad)
    pad3(p, Bf)[..., 1] = 1.0
    J = np.zeros(3 * p.ninner)
    c = (3, 2, 2)                             # (z, y, x), an interior point
    inn3(p, J)[c][0] = 1.0                    # J x B = z^ there only
    k = slope(p, J, Bf)[..., 2]
    want = np.zeros_like(k)
    want[c] = 1 - w
    for d in range(3):
        for s in (-1, 1):
            q = list(c)
            q[d] += s
            want[tuple(q)] = w / 6
    np.testing.assert_allclose(k, want, rtol=1e-15, atol=1e-16)
