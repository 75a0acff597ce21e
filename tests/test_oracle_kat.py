"""Known-answer tests that pin the CPU oracle (no GPU).

The reference ships no golden vectors for the relaxation path (its only
fixture is src/params.in), so each oracle function is pinned by what the
mathematics fixes: the Cash-Karp order conditions, affine maps that make the
central differences exact, Stokes' theorem on planar diamonds, the closed
form of the line-tied face jacobian, the pressure force against automatic
differentiation of beta/det, brute-force reductions and the step controller's
closed form.  Each test is chosen so that a dropped term, a wrong sign or a
transposed index in the oracle fails it.
"""
import ctypes as C

import numpy as np
import pytest

import oracle as o
from glemur_amd import params_default, params_derive, workloads


# ----------------------------------------------------------------- helpers
def grid(nx=9, ny=7, nz=11, L=(4.0, 3.0, 10.0), O=(-2.0, -1.5, -5.0), **kw):
    """Oracle parameters of a box with dyadic spacing (exact grid arithmetic):
    defaults give dx = 0.5, dy = 0.5, dz = 1."""
    p = params_default()
    p.nx, p.ny, p.nz = nx, ny, nz
    p.Lx, p.Ly, p.Lz = L
    p.Ox, p.Oy, p.Oz = O
    p.set_str("bInit", "homZ")
    p.ampl = 1.0
    p.epsilonProf = False        # the reference default is on
    p.pressure = False
    for k, v in kw.items():
        setattr(p, k, v)
    params_derive(p)
    return o.from_product(p)


def pad3(p, a):
    """(nz+2, ny+2, nx+2, 3) view of a padded AoS array."""
    return a.reshape(p.nz + 2, p.ny + 2, p.nx + 2, 3)


def pad1(p, a):
    return a.reshape(p.nz + 2, p.ny + 2, p.nx + 2)


def inn3(p, a):
    """(nz, ny, nx, 3) view of an inner AoS array."""
    return a.reshape(p.nz, p.ny, p.nx, 3)


def lattice(p):
    """Undistorted padded positions X = (i-1) h + O as the reference sets them."""
    k, j, i = np.meshgrid(np.arange(p.nz + 2), np.arange(p.ny + 2), np.arange(p.nx + 2), indexing="ij")
    X = np.stack([p.dx * (i - 1) + p.Ox, p.dy * (j - 1) + p.Oy, p.dz * (k - 1) + p.Oz], axis=-1)
    return X


def empty_pad(p, comps=3):
    return np.zeros(comps * p.npad)


def ref(p):
    return C.byref(p)


# ------------------------------------------------------- a1: tableau
def test_cash_karp_tableau_order_conditions():
    """src/rungeKutta.cu:10-25.  Row sums give the nodes; the 5th-order
    weights satisfy all 17 order conditions up to order 5, the embedded
    weights those up to order 4 and not the order-5 bushy tree."""
    a, b, c, cs = o.coefficients()
    B = b.reshape(7, 6)
    assert np.all(B[0] == 0) and np.all(B[6] == 0)
    for n in range(1, 6):
        assert np.all(B[n, n:] == 0)
        assert abs(B[n, :n].sum() - a[n]) < 1e-15
    A = B[:6, :6]
    e = np.ones(6)
    aa = A @ e

    def conds(w, order):
        t = []
        t += [(w @ e, 1.0)]
        if order >= 2:
            t += [(w @ aa, 1 / 2)]
        if order >= 3:
            t += [(w @ aa**2, 1 / 3), (w @ (A @ aa), 1 / 6)]
        if order >= 4:
            t += [(w @ aa**3, 1 / 4), (w @ (aa * (A @ aa)), 1 / 8), (w @ (A @ aa**2), 1 / 12),
                  (w @ (A @ A @ aa), 1 / 24)]
        if order >= 5:
            t += [(w @ aa**4, 1 / 5), (w @ (aa**2 * (A @ aa)), 1 / 10), (w @ ((A @ aa) ** 2), 1 / 20),
                  (w @ (aa * (A @ aa**2)), 1 / 15), (w @ (A @ aa**3), 1 / 20), (w @ (aa * (A @ A @ aa)), 1 / 30),
                  (w @ (A @ (aa * (A @ aa))), 1 / 40), (w @ (A @ A @ aa**2), 1 / 60),
                  (w @ (A @ A @ A @ aa), 1 / 120)]
        return t

    for got, want in conds(c, 5):
        assert abs(got - want) < 1e-14, (got, want)
    for got, want in conds(cs, 4):
        assert abs(got - want) < 1e-14, (got, want)
    # the embedded pair differs at order 5 (else the error estimate would vanish)
    assert abs(cs @ aa**4 - 1 / 5) > 1e-4


def test_stage_chain_round_trip():
    """src/rungeKutta.cu:216-227: x_{n+1} = x_n + sum_{m<=n}(b[n+1][m]-b[n][m]) k_m
    telescopes to x_0 + sum_m b[n+1][m] k_m, and returns to x_0 after stage 5."""
    p = grid(nx=5, ny=4, nz=6)
    a, b, c, cs = o.coefficients()
    rng = np.random.default_rng(1)
    x0 = rng.standard_normal(3 * p.npad)
    Bf = rng.standard_normal(3 * p.npad)
    G = np.zeros(3 * p.ninner)
    kk = np.zeros(18 * p.ninner)
    xin = x0.copy()
    dt = 0.37
    Bm = b.reshape(7, 6)
    ks = []
    for n in range(6):
        J = rng.standard_normal(3 * p.ninner)
        xt = xin.copy()
        o.lib.or_kk(ref(p), o.dp(b), n, dt, o.dp(xin), o.dp(Bf), o.dp(J), o.dp(G), o.dp(kk), o.dp(xt))
        kn = inn3(p, kk[3 * p.ninner * n: 3 * p.ninner * (n + 1)])
        want_k = np.cross(inn3(p, J), pad3(p, Bf)[1:-1, 1:-1, 1:-1]) * dt
        np.testing.assert_allclose(kn, want_k, rtol=1e-15, atol=1e-15)
        ks.append(kn.copy())
        want_x = pad3(p, x0)[1:-1, 1:-1, 1:-1] + sum(Bm[n + 1, m] * ks[m] for m in range(n + 1))
        np.testing.assert_allclose(pad3(p, xt)[1:-1, 1:-1, 1:-1], want_x, rtol=0, atol=1e-14)
        # ghosts untouched
        g = ~_inner(p)
        assert np.array_equal(pad3(p, xt)[g], pad3(p, xin)[g])
        xin = xt
    np.testing.assert_allclose(pad3(p, xin), pad3(p, x0), rtol=0, atol=1e-14)


def _inner(p):
    m = np.zeros((p.nz + 2, p.ny + 2, p.nx + 2), dtype=bool)
    m[1:-1, 1:-1, 1:-1] = True
    return m


def test_kk_pressure_term():
    """k = dt (J x B) eps - dt gradP (src/rungeKutta.cu:169-193)."""
    p = grid(nx=4, ny=3, nz=5, pressure=True)
    _, b, _, _ = o.coefficients()
    rng = np.random.default_rng(2)
    xb = rng.standard_normal(3 * p.npad)
    Bf = rng.standard_normal(3 * p.npad)
    J = rng.standard_normal(3 * p.ninner)
    G = rng.standard_normal(3 * p.ninner)
    kk = np.zeros(18 * p.ninner)
    xt = xb.copy()
    dt = 0.125
    o.lib.or_kk(ref(p), o.dp(b), 0, dt, o.dp(xb), o.dp(Bf), o.dp(J), o.dp(G), o.dp(kk), o.dp(xt))
    want = np.cross(inn3(p, J), pad3(p, Bf)[1:-1, 1:-1, 1:-1]) * dt - dt * inn3(p, G)
    np.testing.assert_allclose(inn3(p, kk[:3 * p.ninner]), want, rtol=1e-14, atol=1e-15)


@pytest.mark.parametrize("n", [(8, 9, 10), (7, 12, 5)])
def test_epsilon_profile(n):
    """eps = prod_d (1 - 4 (i - n/2)(i + 1 - n/2) / n^2) (src/rungeKutta.cu:180-185):
    2/n at both ends of each axis, symmetric, quadratic, 1 at the centre of
    an even axis."""
    nx, ny, nz = n
    p = grid(nx=nx, ny=ny, nz=nz, epsilonProf=True)
    _, b, _, _ = o.coefficients()
    xb = empty_pad(p)
    Bf = np.zeros(3 * p.npad)
    pad3(p, Bf)[..., 1] = 1.0                 # B = y^
    J = np.zeros(3 * p.ninner)
    inn3(p, J)[..., 0] = 1.0                  # J = x^  ->  J x B = z^
    G = np.zeros(3 * p.ninner)
    kk = np.zeros(18 * p.ninner)
    xt = xb.copy()
    o.lib.or_kk(ref(p), o.dp(b), 0, 1.0, o.dp(xb), o.dp(Bf), o.dp(J), o.dp(G), o.dp(kk), o.dp(xt))
    eps = inn3(p, kk[:3 * p.ninner])[..., 2]   # (nz, ny, nx)
    # separable: eps = ex(i) ey(j) ez(k)
    ex = eps[0, 0, :] / eps[0, 0, 0] * (2 / nx)
    ey = eps[0, :, 0] / eps[0, 0, 0] * (2 / ny)
    ez = eps[:, 0, 0] / eps[0, 0, 0] * (2 / nz)
    assert abs(eps[0, 0, 0] - (2 / nx) * (2 / ny) * (2 / nz)) < 1e-15
    np.testing.assert_allclose(eps, ez[:, None, None] * ey[None, :, None] * ex[None, None, :], rtol=1e-13)
    for e, m in ((ex, nx), (ey, ny), (ez, nz)):
        np.testing.assert_allclose(e, e[::-1], rtol=1e-13)
        assert abs(e[-1] - 2 / m) < 1e-14
        if m >= 4:
            np.testing.assert_allclose(np.diff(e, 3), 0, atol=1e-13)
        if m % 2 == 0:
            assert abs(e[m // 2] - 1) < 1e-14 and abs(e[m // 2 - 1] - 1) < 1e-14


# ------------------------------------------------------- a2: B_JacB0
def test_b_jacb0_identity_bitwise():
    """Undistorted dyadic grid: jac = I exactly, so B = B0 and det = 1 bitwise."""
    p = grid()
    rng = np.random.default_rng(3)
    xb = lattice(p).reshape(-1).copy()
    B0 = rng.standard_normal(3 * p.npad)
    B = np.zeros(3 * p.npad)
    det = np.zeros(p.npad)
    o.lib.or_b_jacb0(ref(p), o.dp(xb), o.dp(B0), o.dp(B), o.dp(det))
    m = _inner(p)
    assert np.array_equal(pad3(p, B)[m], pad3(p, B0)[m])
    assert np.all(pad1(p, det)[m] == 1.0)


def test_b_jacb0_affine():
    """x = A X + c: central differences are exact, so det = det A and
    B = A B0 / det A (src/rungeKutta.cu:28-108)."""
    p = grid()
    rng = np.random.default_rng(4)
    A = np.eye(3) + 0.3 * rng.standard_normal((3, 3))
    assert np.linalg.det(A) > 0.2
    X = lattice(p)
    xb = (X @ A.T + np.array([0.3, -1.1, 2.0])).reshape(-1).copy()
    B0 = rng.standard_normal(3 * p.npad)
    B = np.zeros(3 * p.npad)
    det = np.zeros(p.npad)
    o.lib.or_b_jacb0(ref(p), o.dp(xb), o.dp(B0), o.dp(B), o.dp(det))
    m = _inner(p)
    dA = np.linalg.det(A)
    np.testing.assert_allclose(pad1(p, det)[m], dA, rtol=1e-13)
    np.testing.assert_allclose(pad3(p, B)[m], pad3(p, B0)[m] @ A.T / dA, rtol=1e-12, atol=1e-13)


# ------------------------------------------------------- a3: updateBbound
def _face_points(p):
    """(face, side, k, j, i) of every point with exactly one ghost coordinate."""
    pts = []
    nx, ny, nz = p.nx, p.ny, p.nz
    for side in (0, 1):
        kg = nz + 1 if side else 0
        for j in range(1, ny + 1):
            for i in range(1, nx + 1):
                pts.append((2, side, kg, j, i))
        jg = ny + 1 if side else 0
        for k in range(1, nz + 1):
            for i in range(1, nx + 1):
                pts.append((1, side, k, jg, i))
        ig = nx + 1 if side else 0
        for k in range(1, nz + 1):
            for j in range(1, ny + 1):
                pts.append((0, side, k, j, ig))
    return pts


def test_bbound_undistorted_bitwise():
    """Undistorted dyadic grid: every face point gets B = B0, det = 1."""
    p = grid()
    rng = np.random.default_rng(5)
    xb = lattice(p).reshape(-1).copy()
    B0 = rng.standard_normal(3 * p.npad)
    B = np.full(3 * p.npad, 7.0)
    det = np.full(p.npad, 7.0)
    o.lib.or_update_bbound(ref(p), o.dp(xb), o.dp(B0), o.dp(B), o.dp(det))
    for f, s, k, j, i in _face_points(p):
        assert np.array_equal(pad3(p, B)[k, j, i], pad3(p, B0)[k, j, i])
        assert pad1(p, det)[k, j, i] == 1.0


@pytest.mark.parametrize("dims", [(9, 7, 11), (20, 18, 5)])
def test_bbound_line_tied_closed_form(dims):
    """Line-tied faces with a moved interior (src/boundaries.cu:39-446): the
    boundary layer is the undistorted lattice, the outer ghost sits one h
    further out along the normal, so the face jacobian has unit in-plane
    columns and the normal column v = (x_in - x_out) / 2h.  Then det = v_n,
    the normal component of B is B0_n and the tangential ones are
    (B0_t + v_t B0_n) / v_n.  Grids wider than 16 cross the reference's
    launch-tile edges, where in-plane neighbours are rebuilt as x +- h."""
    nx, ny, nz = dims
    p = grid(nx=nx, ny=ny, nz=nz, L=(0.25 * (nx - 1), 0.25 * (ny - 1), 0.5 * (nz - 1)))
    rng = np.random.default_rng(6)
    X = lattice(p)
    x = X.copy()
    x[1:-1, 1:-1, 1:-1] += 0.05 * rng.standard_normal(x[1:-1, 1:-1, 1:-1].shape)
    xb = x.reshape(-1).copy()
    B0 = rng.standard_normal(3 * p.npad)
    B = np.zeros(3 * p.npad)
    det = np.zeros(p.npad)
    o.lib.or_update_bbound(ref(p), o.dp(xb), o.dp(B0), o.dp(B), o.dp(det))
    h = (p.dx, p.dy, p.dz)
    n_checked = 0
    for f, s, k, j, i in _face_points(p):
        idx = [i, j, k]
        inn = list(idx)
        inn[f] += -1 if s else 1
        e = np.zeros(3)
        e[f] = 1.0
        xg = x[k, j, i]
        xi = x[inn[2], inn[1], inn[0]]
        xo = xg + (e * h[f] if s else -e * h[f])
        v = (xo - xi) / (2 * h[f]) if s else (xi - xo) / (2 * h[f])
        b0 = pad3(p, B0)[k, j, i]
        want = (b0 + v * b0[f]) / v[f]
        want[f] = b0[f]
        got = pad3(p, B)[k, j, i]
        np.testing.assert_allclose(got, want, rtol=1e-13, atol=1e-13)
        assert got[f] == b0[f]                      # B.n = B0.n exactly
        assert abs(pad1(p, det)[k, j, i] - v[f]) < 1e-13 * abs(v[f])
        n_checked += 1
    assert n_checked == 2 * (nx * ny + ny * nz + nx * nz)


def test_bbound_leaves_edges_and_corners():
    """Ghost edges and corners are never written (B there stays B0)."""
    p = grid()
    rng = np.random.default_rng(7)
    x = lattice(p)
    x[1:-1, 1:-1, 1:-1] += 0.05 * rng.standard_normal(x[1:-1, 1:-1, 1:-1].shape)
    xb = x.reshape(-1).copy()
    B0 = rng.standard_normal(3 * p.npad)
    B = np.full(3 * p.npad, 3.5)
    det = np.full(p.npad, 3.5)
    o.lib.or_update_bbound(ref(p), o.dp(xb), o.dp(B0), o.dp(B), o.dp(det))
    k, j, i = np.meshgrid(np.arange(p.nz + 2), np.arange(p.ny + 2), np.arange(p.nx + 2), indexing="ij")
    ng = ((i == 0) | (i == p.nx + 1)).astype(int) + ((j == 0) | (j == p.ny + 1)) + ((k == 0) | (k == p.nz + 1))
    keep = (ng != 1)
    assert np.all(pad3(p, B)[keep] == 3.5) and np.all(pad1(p, det)[keep] == 3.5)


# ------------------------------------------------------- a4: periodic copies
def test_periodic_copies():
    """setPeri copies the opposite inner plane; setGridPeri adds -+(L + h)
    along the periodic axis only (src/boundaries.cu:450-582)."""
    p = grid(xPeri=True, yPeri=True, zPeri=True)
    rng = np.random.default_rng(8)
    f = rng.standard_normal(3 * p.npad)
    g = f.copy()
    o.lib.or_set_peri(ref(p), o.dp(g))
    F, G = pad3(p, f), pad3(p, g)
    nx, ny, nz = p.nx, p.ny, p.nz
    # x applied last: x ghost planes copy the (already y/z-updated) planes
    np.testing.assert_array_equal(G[:, :, 0], G[:, :, nx])
    np.testing.assert_array_equal(G[:, :, nx + 1], G[:, :, 1])
    np.testing.assert_array_equal(G[1:-1, 0, 1:-1], F[1:-1, ny, 1:-1])
    np.testing.assert_array_equal(G[0, 1:-1, 1:-1], F[nz, 1:-1, 1:-1])
    np.testing.assert_array_equal(G[1:-1, 1:-1, 1:-1], F[1:-1, 1:-1, 1:-1])

    x = lattice(p).reshape(-1).copy()
    x2 = x.copy()
    o.lib.or_set_grid_peri(ref(p), o.dp(x2))
    X0, X2 = pad3(p, x), pad3(p, x2)
    # on the undistorted lattice the shifted copy reproduces the lattice
    np.testing.assert_allclose(X2, X0, rtol=0, atol=1e-13)
    # and on a moved grid the shift is exactly -+(L+h) on the axis
    xm = x.copy()
    pad3(p, xm)[1:-1, 1:-1, 1:-1] += 0.01 * rng.standard_normal((nz, ny, nx, 3))
    xm2 = xm.copy()
    o.lib.or_set_grid_peri(ref(p), o.dp(xm2))
    M, M2 = pad3(p, xm), pad3(p, xm2)
    np.testing.assert_allclose(M2[1:-1, 1:-1, 0, 0], M[1:-1, 1:-1, nx, 0] - (p.Lx + p.dx), rtol=0, atol=1e-14)
    np.testing.assert_array_equal(M2[1:-1, 1:-1, 0, 1:], M[1:-1, 1:-1, nx, 1:])
    np.testing.assert_allclose(M2[1:-1, ny + 1, 1:-1, 1], M[1:-1, 1, 1:-1, 1] + (p.Ly + p.dy), rtol=0, atol=1e-14)


# ------------------------------------------------------- a5: JStokes
@pytest.mark.parametrize("seed", [10, 11])
def test_jstokes_linear_field_affine_grid(seed):
    """Linear B = M x + b0 on an affine grid: the diamonds are planar
    parallelograms with straight edges, so the trapezoid circulation and the
    area vectors are exact and Stokes gives J = curl B
    = (M21 - M12, M02 - M20, M10 - M01) at every inner point
    (src/current.cu:222-409)."""
    p = grid(nx=6, ny=5, nz=7)
    rng = np.random.default_rng(seed)
    A = np.eye(3) + 0.25 * rng.standard_normal((3, 3))
    x = lattice(p) @ A.T + rng.standard_normal(3)
    M = rng.standard_normal((3, 3))
    Bf = x @ M.T + rng.standard_normal(3)
    J = np.zeros(3 * p.ninner)
    o.lib.or_jstokes(ref(p), o.dp(x.reshape(-1).copy()), o.dp(Bf.reshape(-1).copy()), o.dp(J))
    curl = np.array([M[2, 1] - M[1, 2], M[0, 2] - M[2, 0], M[1, 0] - M[0, 1]])
    np.testing.assert_allclose(inn3(p, J), np.broadcast_to(curl, (p.nz, p.ny, p.nx, 3)), rtol=1e-10, atol=1e-10)


def test_jstokes_equals_cofactor_form():
    """On any grid the reference's diamond circulation over area equals
    J = (C_x x db_x + C_y x db_y + C_z x db_z) / det[dx dy dz] with d. the
    central differences x(+a) - x(-a), B(+a) - B(-a) and C_x = dy x dz,
    C_y = dz x dx, C_z = dx x dy (exact algebra: the trapezoid circulation of
    a diamond is (db02 . dr13 - db13 . dr02) / 2, its mean area dr02 x dr13 / 2,
    and the 3x3 system has the cofactor vectors as rows).  Random distorted
    grid and random B, numpy vs the oracle's reference-order sums."""
    p = grid(nx=7, ny=6, nz=8)
    rng = np.random.default_rng(16)
    x = lattice(p)
    x[1:-1, 1:-1, 1:-1] += 0.08 * rng.standard_normal(x[1:-1, 1:-1, 1:-1].shape)
    Bf = rng.standard_normal(x.shape)
    J = np.zeros(3 * p.ninner)
    o.lib.or_jstokes(ref(p), o.dp(x.reshape(-1).copy()), o.dp(Bf.reshape(-1).copy()), o.dp(J))
    c = (slice(1, -1),) * 3

    def d(f, ax):
        sp = [slice(1, -1)] * 3
        sm = [slice(1, -1)] * 3
        sp[2 - ax] = slice(2, None)
        sm[2 - ax] = slice(0, -2)
        return f[tuple(sp)] - f[tuple(sm)]

    dr = [d(x, a) for a in range(3)]
    db = [d(Bf, a) for a in range(3)]
    C = [np.cross(dr[1], dr[2]), np.cross(dr[2], dr[0]), np.cross(dr[0], dr[1])]
    D = np.einsum("...l,...l->...", dr[0], C[0])
    want = sum(np.cross(C[a], db[a]) for a in range(3)) / D[..., None]
    np.testing.assert_allclose(inn3(p, J), want, rtol=1e-10, atol=1e-10 * np.abs(want).max())


def test_jstokes_uniform_field_zero():
    p = grid(nx=5, ny=6, nz=4)
    rng = np.random.default_rng(12)
    x = lattice(p)
    x[1:-1, 1:-1, 1:-1] += 0.05 * rng.standard_normal(x[1:-1, 1:-1, 1:-1].shape)
    Bf = np.broadcast_to(rng.standard_normal(3), x.shape).copy()
    J = np.zeros(3 * p.ninner)
    o.lib.or_jstokes(ref(p), o.dp(x.reshape(-1).copy()), o.dp(Bf.reshape(-1)), o.dp(J))
    np.testing.assert_allclose(J, 0, atol=1e-12)


# ------------------------------------------------------- a6: gradP
def test_gradp_affine_zero():
    """p = beta/det is uniform under an affine map: gradP = 0."""
    p = grid(beta=0.7, pressure=True)
    rng = np.random.default_rng(13)
    A = np.eye(3) + 0.2 * rng.standard_normal((3, 3))
    x = (lattice(p) @ A.T).reshape(-1).copy()
    B0 = np.zeros(3 * p.npad)
    B = np.zeros(3 * p.npad)
    det = np.zeros(p.npad)
    o.lib.or_b_jacb0(ref(p), o.dp(x), o.dp(B0), o.dp(B), o.dp(det))
    G = np.zeros(3 * p.ninner)
    o.lib.or_gradp(ref(p), o.dp(x), o.dp(det), o.dp(G))
    np.testing.assert_allclose(G, 0, atol=1e-11)


def _smooth_map_torch(Xt, alpha, k):
    import torch
    X, Y, Z = Xt[..., 0], Xt[..., 1], Xt[..., 2]
    return torch.stack([X + alpha * torch.sin(k * Y) * torch.sin(k * Z) + 0.5 * alpha * torch.sin(k * X),
                        Y + alpha * torch.sin(k * Z) * torch.cos(k * X),
                        Z + alpha * torch.sin(k * X) * torch.sin(k * Y) + 0.3 * alpha * torch.cos(k * Z)], -1)


def _gradp_exact(points, beta, alpha, k):
    """grad_x (beta / det(dx/dX)) by automatic differentiation, an
    independent route from the oracle's second-difference formula."""
    import torch
    from torch.func import jacrev, vmap, grad

    def fmap(X):
        return _smooth_map_torch(X, alpha, k)

    def pfun(X):
        return beta / torch.linalg.det(jacrev(fmap)(X))

    Xt = torch.tensor(points, dtype=torch.float64)
    gX = vmap(grad(pfun))(Xt)                    # grad_X p = J^T grad_x p
    Jm = vmap(jacrev(fmap))(Xt)
    gx = torch.linalg.solve(Jm.transpose(-1, -2), gX.unsqueeze(-1)).squeeze(-1)
    return gx.numpy()


def test_gradp_converges_to_grad_beta_over_det():
    """gradP (src/gradP.cu:10-156) is beta * grad_x(1/det): compared with
    automatic differentiation of a smooth 3D map, the error falls like h^2."""
    import torch
    beta, alpha, kw = 0.8, 0.08, 1.3
    errs = []
    for n in (9, 17):
        p = grid(nx=n, ny=n, nz=n, L=(2.0, 2.0, 2.0), O=(-1.0, -1.0, -1.0), beta=beta, pressure=True)
        X = lattice(p)
        with torch.no_grad():
            x = _smooth_map_torch(torch.tensor(X), alpha, kw).numpy()
        xb = x.reshape(-1).copy()
        B0 = np.zeros(3 * p.npad)
        B = np.zeros(3 * p.npad)
        det = np.zeros(p.npad)
        o.lib.or_b_jacb0(ref(p), o.dp(xb), o.dp(B0), o.dp(B), o.dp(det))
        G = np.zeros(3 * p.ninner)
        o.lib.or_gradp(ref(p), o.dp(xb), o.dp(det), o.dp(G))
        want = _gradp_exact(X[1:-1, 1:-1, 1:-1].reshape(-1, 3), beta, alpha, kw)
        err = np.abs(G.reshape(-1, 3) - want).max()
        errs.append(err)
        assert np.abs(want).max() > 0.05
    assert errs[1] < errs[0] / 3.2, errs
    assert errs[1] < 5e-3, errs


# ------------------------------------------------------- a8-a12: reductions
def _weights(p):
    w = [np.ones(m) for m in (p.nx, p.ny, p.nz)]
    for a in w:
        a[0] *= 0.5
        a[-1] *= 0.5
    return w[2][:, None, None] * w[1][None, :, None] * w[0][None, None, :]


def test_b2_uniform_field_closed_form():
    """Undistorted grid, B = ampl z^: B2 = Lx Ly Lz ampl^2 (the 1/2^bound
    weights make it the trapezoid rule; src/maths.cu:647-680)."""
    p = grid(nx=9, ny=7, nz=11)
    ampl = 1.7
    Bf = np.zeros(3 * p.npad)
    pad3(p, Bf)[..., 2] = ampl
    det = np.ones(p.npad)
    got = o.lib.or_b2(ref(p), o.dp(Bf), o.dp(det))
    assert abs(got - p.Lx * p.Ly * p.Lz * ampl ** 2) < 1e-12 * got


def test_reductions_brute_force():
    p = grid(nx=5, ny=6, nz=7)
    rng = np.random.default_rng(14)
    Bf = rng.standard_normal(3 * p.npad)
    det = 0.5 + rng.random(p.npad)
    J = rng.standard_normal(3 * p.ninner)
    Ji = inn3(p, J)
    Ji[0, 0, 0] = 0.0                      # |J||B| = 0 there: ratio 0 by definition
    Ji[1, 2, 3] = np.nan                   # NaN is skipped by the max
    Bi = pad3(p, Bf)[1:-1, 1:-1, 1:-1]
    di = pad1(p, det)[1:-1, 1:-1, 1:-1]
    b2 = (Bi ** 2).sum(-1)
    want_b2 = (b2 * di * _weights(p)).sum() * p.dx * p.dy * p.dz
    assert abs(o.lib.or_b2(ref(p), o.dp(Bf), o.dp(det)) - want_b2) < 1e-12 * abs(want_b2)
    nj = np.sqrt((Ji ** 2).sum(-1))
    assert o.lib.or_jmax(ref(p), o.dp(J)) == np.nanmax(nj)
    c = np.cross(Ji, Bi)
    nc = np.sqrt((c ** 2).sum(-1))
    r1 = nc / b2
    assert o.lib.or_jxb_b2_max(ref(p), o.dp(Bf), o.dp(J)) == pytest.approx(np.nanmax(r1), rel=1e-15)
    with np.errstate(invalid="ignore", divide="ignore"):
        r2 = np.where(nj * np.sqrt(b2) > 0, nc / (nj * np.sqrt(b2)), 0.0)
    assert o.lib.or_jxb_jb_max(ref(p), o.dp(Bf), o.dp(J)) == pytest.approx(np.nanmax(r2), rel=1e-15)
    assert np.nanmax(r2) <= 1.0 + 1e-15
    want = np.nanmax(nc) / np.nanmax(nj * np.sqrt(b2))
    assert o.lib.or_jxbmax_jbmax(ref(p), o.dp(Bf), o.dp(J)) == pytest.approx(want, rel=1e-15)
    # force-free: J parallel to B everywhere gives 0; J = 0 gives 0 (no 0/0)
    Jff = (Bi * 0.7).reshape(-1).copy()
    assert o.lib.or_jxbmax_jbmax(ref(p), o.dp(Bf), o.dp(Jff)) < 1e-15
    Jz = np.zeros_like(J)
    assert o.lib.or_jxbmax_jbmax(ref(p), o.dp(Bf), o.dp(Jz)) == 0.0


@pytest.mark.parametrize("zupdate", [True, False])
def test_xnewstar_brute_force(zupdate):
    """x5 = x + sum c_m k_m, x4 = x + sum c*_m k_m (z only with zUpdate);
    the max of |x5 - x4| skips NaN and flags it (src/rungeKutta.cu:251-353,
    src/maths.cu:144-153)."""
    p = grid(nx=4, ny=5, nz=3, zUpdate=zupdate)
    _, _, c, cs = o.coefficients()
    rng = np.random.default_rng(15)
    xb = rng.standard_normal(3 * p.npad)
    kk = rng.standard_normal(18 * p.ninner) * 1e-3
    xn = np.zeros(3 * p.npad)
    nan = C.c_int(0)
    md = o.lib.or_xnewstar(ref(p), o.dp(c), o.dp(cs), o.dp(xb), o.dp(kk), o.dp(xn), C.byref(nan))
    K = kk.reshape(6, p.nz, p.ny, p.nx, 3)
    x = pad3(p, xb)[1:-1, 1:-1, 1:-1]
    x5 = x + np.tensordot(c, K, axes=1)
    x4 = x + np.tensordot(cs, K, axes=1)
    if not zupdate:
        x5[..., 2] = x[..., 2]
        x4[..., 2] = x[..., 2]
    np.testing.assert_allclose(pad3(p, xn)[1:-1, 1:-1, 1:-1], x5, rtol=0, atol=1e-15)
    assert md == pytest.approx(np.abs(x5 - x4).max(), rel=1e-9)
    assert nan.value == 0
    K[2, 1, 1, 1, 0] = np.nan
    md2 = o.lib.or_xnewstar(ref(p), o.dp(c), o.dp(cs), o.dp(xb), o.dp(kk), o.dp(xn), C.byref(nan))
    assert nan.value == 1 and np.isfinite(md2) and md2 > 0


# ------------------------------------------------------- a10: controller
def test_step_controller_closed_form():
    """Accept iff maxDelta < maxError; dt <- dt (0.9 maxError/maxDelta)^0.2
    after accept and reject; a reject leaves x at the stage-6 position
    (= x0 up to round-off) and t unchanged (src/main.cu:318-383)."""
    from tests.conftest import small_braid
    pp = small_braid(9, 10, 12)
    p = o.from_product(pp)
    B0, xb = o.init_state(p)
    sim = o.Sim(p, B0, xb)
    seen = set()
    for it in range(14):
        if it == 6:
            sim.set_dt(50 * sim.dt)          # force a reject
        x_before = sim.arr("xb").copy()
        t0, dt0 = sim.t, sim.dt
        acc, dt_used, md, nan = sim.step()
        assert dt_used == dt0 and not nan
        assert acc == (md < p.maxError)
        assert sim.dt == pytest.approx(dt0 * (0.9 * p.maxError / md) ** 0.2, rel=1e-15)
        if acc:
            assert sim.t == t0 + dt0
        else:
            assert sim.t == t0
            # round-off of the telescoped stage chain scales with the slopes
            kmax = np.abs(sim.arr("kk")).max()
            np.testing.assert_allclose(sim.arr("xb"), x_before, rtol=0, atol=1e-14 * max(1.0, kmax))
        seen.add(acc)
    assert seen == {True, False}


def test_energy_decreases_under_magnetofriction():
    """dE/dt = -int |J x B|^2 dV <= 0: the oracle's B2 falls on every accepted
    step of the braid relaxation."""
    from tests.conftest import small_braid
    p = o.from_product(small_braid(11, 11, 15))
    B0, xb = o.init_state(p)
    sim = o.Sim(p, B0, xb)
    sim.fields()
    e_prev = sim.b2()
    n_acc = 0
    for _ in range(25):
        acc, *_ = sim.step()
        if acc:
            e = sim.b2()
            assert e < e_prev
            e_prev = e
            n_acc += 1
    assert n_acc >= 15


# ------------------------------------------------------- a14: init
def test_init_blobs_closed_form():
    """B0 = ampl z^ + sum_b 2 ampl twist_b/scale_b exp(-rho_b^2/scale_b^2 -
    (z - zc_b)^2/zl_b^2) (-(y - yc_b), x - xc_b, 0); x = (i-1) h + O over
    the padded indices, so the inner points span [O, O + L]
    (src/init.cu:109-124, 223-226)."""
    pp = workloads.braid_params(9, 11, 13)
    p = o.from_product(pp)
    B0, xb = o.init_state(p)
    X = lattice(p)
    np.testing.assert_array_equal(pad3(p, xb), X)
    x = pad3(p, xb)
    assert x[1, 1, 1, 0] == p.Ox and abs(x[1, 1, p.nx, 0] - (p.Ox + p.Lx)) < 1e-14
    assert abs(x[p.nz, 1, 1, 2] - (p.Oz + p.Lz)) < 1e-13
    want = np.zeros_like(X)
    want[..., 2] = p.ampl
    for m in range(p.nBlobs):
        dxc = X[..., 0] - p.blobXc[m]
        dyc = X[..., 1] - p.blobYc[m]
        dzc = X[..., 2] - p.blobZc[m]
        e = np.exp(-(dxc ** 2 + dyc ** 2) / p.blobScale[m] ** 2 - dzc ** 2 / p.blobZl[m] ** 2)
        cm = 2 * p.ampl * p.blobTwist[m] / p.blobScale[m]
        assert abs(abs(cm) - 1.0) < 1e-15           # c_k = +-1 for the braid
        want[..., 0] += -cm * e * dyc
        want[..., 1] += cm * e * dxc
    np.testing.assert_allclose(pad3(p, B0), want, rtol=1e-13, atol=1e-15)


def test_init_residual():
    """initResiduals (src/init.cu:13-26): Lx Ly Lz ampl for the closed-form fields."""
    p = o.from_product(workloads.braid_params(9, 9, 9))
    assert o.lib.or_init_residual(ref(p)) == p.Lx * p.Ly * p.Lz * p.ampl
