"""The inertia equation of the oracle (src/rungeKutta.cu:112-353, inertia
branches; U2/UMax of src/diagnostics.cu:114-130), pinned by closed forms:

- at rest the velocity slope is det times the magnetofrictional Lorentz
  slope plus the same pressure term, and the position slope vanishes;
- a uniformly compressed grid x = s X with B0 = ez has B = ez / s^2,
  det = s^3 and J = 0 away from the faces, so the velocity slope there is
  -dt nu s^3 u for any uniform u;
- an all-periodic grid carried by a uniform velocity stays a rigid
  translation (B uniform, J = 0, det = 1), so u' = -nu u and x' = u give
  u(t) = u0 exp(-nu t), x(t) = x0 + u0 (1 - exp(-nu t)) / nu; the adaptive
  Cash-Karp loop must follow both through accepted and rejected steps;
- U2 of a uniform velocity on an undistorted grid is |u|^2 Lx Ly Lz (the
  boundary weights make the trapezoid rule), UMax the largest |u| with NaN
  skipped."""
import numpy as np

import oracle as o
from test_oracle_kat import grid, inn3, lattice, pad3, ref


def fields(p, B0, xb):
    """B, det, J, gradP of a state through the oracle's per-kernel calls."""
    sim = o.Sim(p, B0, xb)
    sim.fields()
    if p.pressure:
        sim.gradp()
    return sim


def stage0(p, sim, dt, uu, inertia):
    """kk[0] of the magnetofrictional or the inertia branch."""
    b = o.coefficients()[1]
    N = p.ninner
    xt = sim.arr("xb").copy()
    if inertia:
        kk = np.zeros(36 * N)
        ut = np.zeros(3 * N)
        o.lib.or_kk_inertia(ref(p), o.dp(b), 0, dt, sim.s.xb, sim.s.B, sim.s.J, sim.s.gradP, sim.s.detJac,
                            o.dp(uu), o.dp(kk), o.dp(xt), o.dp(ut))
        k = kk[:6 * N].reshape(N, 6)
        return k[:, :3], k[:, 3:], xt, ut
    kk = np.zeros(18 * N)
    o.lib.or_kk(ref(p), o.dp(b), 0, dt, sim.s.xb, sim.s.B, sim.s.J, sim.s.gradP, o.dp(kk), o.dp(xt))
    return kk[:3 * N].reshape(N, 3), xt


def distorted(p, seed):
    rng = np.random.default_rng(seed)
    X = lattice(p)
    x = X.copy()
    h = np.array([p.dx, p.dy, p.dz])
    x[1:-1, 1:-1, 1:-1] += 0.08 * h * rng.standard_normal((p.nz, p.ny, p.nx, 3))
    B0 = np.zeros_like(X)
    B0[..., 0] = 0.3 * np.sin(X[..., 1])
    B0[..., 1] = 0.2 * np.cos(X[..., 2])
    B0[..., 2] = 1.0
    return B0.reshape(-1).copy(), x.reshape(-1).copy()


def test_slope_at_rest_is_det_times_the_lorentz_slope():
    dt, nu, beta = 0.013, 0.7, 0.05
    B0, xb = distorted(grid(), 31)
    pm = grid(pressure=True, beta=beta)
    pn = grid()
    pi = grid(pressure=True, beta=beta, inertia=True, nu=nu)
    sp, sn, si = fields(pm, B0, xb), fields(pn, B0, xb), fields(pi, B0, xb)
    k_p, _ = stage0(pm, sp, dt, None, False)          # dt JxB - dt gradP
    k_n, _ = stage0(pn, sn, dt, None, False)          # dt JxB
    det = si.arr("detJac").reshape(pi.nz + 2, pi.ny + 2, pi.nx + 2)[1:-1, 1:-1, 1:-1].reshape(-1, 1)
    assert np.ptp(det) > 0.05                         # a distorted grid
    u0 = np.zeros(3 * pi.ninner)
    kx, ku, xt, ut = stage0(pi, si, dt, u0, True)
    assert np.all(kx == 0)
    np.testing.assert_array_equal(xt, si.arr("xb"))           # the positions wait for the velocity
    np.testing.assert_array_equal(ut, 0.2 * ku.reshape(-1))   # u1 = u0 + b10 k_u0, b10 = 1/5
    want = det * k_n - (k_n - k_p)
    np.testing.assert_allclose(ku, want, rtol=1e-12, atol=1e-14 * np.abs(want).max())
    # a moving grid adds the friction term and moves the positions
    rng = np.random.default_rng(5)
    u = rng.standard_normal(3 * pi.ninner)
    kx2, ku2, xt2, ut2 = stage0(pi, si, dt, u, True)
    U = u.reshape(-1, 3)
    np.testing.assert_array_equal(kx2, dt * U)
    np.testing.assert_allclose(ku2 - ku, -dt * nu * U * det, rtol=1e-11, atol=1e-15)
    # stage 1 position and velocity: x + b10 k_0, u + b10 k_u0 with b10 = 1/5
    np.testing.assert_allclose(ut2.reshape(-1, 3), U + 0.2 * ku2, rtol=1e-15, atol=1e-15)
    xi = pad3(pi, xt2)[1:-1, 1:-1, 1:-1].reshape(-1, 3)
    x0 = pad3(pi, si.arr("xb"))[1:-1, 1:-1, 1:-1].reshape(-1, 3)
    np.testing.assert_allclose(xi, x0 + 0.2 * dt * U, rtol=1e-15, atol=1e-15)


def test_compressed_grid_damps_with_det_cubed():
    s, dt, nu = 0.8, 0.02, 1.3
    p = grid(inertia=True, nu=nu)
    X = lattice(p)
    x = (s * X).reshape(-1).copy()
    B0 = np.zeros_like(X)
    B0[..., 2] = 1.0
    sim = fields(p, B0.reshape(-1).copy(), x)
    det = sim.arr("detJac").reshape(p.nz + 2, p.ny + 2, p.nx + 2)[1:-1, 1:-1, 1:-1]
    np.testing.assert_allclose(det, s ** 3, rtol=1e-14)
    Bi = pad3(p, sim.arr("B"))[1:-1, 1:-1, 1:-1]
    np.testing.assert_allclose(Bi, np.broadcast_to([0, 0, s ** -2], Bi.shape), rtol=0, atol=1e-14)
    # J vanishes away from the line-tied faces (their ghost B is built by
    # setBbound, not from the scaled ghost positions)
    core = (slice(1, -1), slice(1, -1), slice(1, -1))
    assert np.abs(inn3(p, sim.arr("J"))[core]).max() < 1e-12
    u0 = np.array([0.3, -0.2, 0.45])
    uu = np.tile(u0, p.ninner)
    kx, ku, _, _ = stage0(p, sim, dt, uu, True)
    ku = ku.reshape(p.nz, p.ny, p.nx, 3)[core]
    np.testing.assert_allclose(ku, np.broadcast_to(-dt * nu * s ** 3 * u0, ku.shape), rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(kx, np.broadcast_to(dt * u0, kx.shape), rtol=1e-15)


def test_periodic_translation_follows_the_damped_motion():
    nu = 0.9
    p = grid(nx=8, ny=6, nz=10, xPeri=True, yPeri=True, zPeri=True, inertia=True, nu=nu, maxError=1e-10,
             dt0=0.8)
    x = lattice(p).reshape(-1).copy()
    o.lib.or_set_grid_peri(ref(p), o.dp(x))
    B0 = np.zeros(3 * p.npad)
    pad3(p, B0)[..., 2] = 1.0
    sim = o.Sim(p, B0, x)
    u0 = np.array([0.21, -0.13, 0.07])
    sim.arr("uu")[:] = np.tile(u0, p.ninner)
    x0 = pad3(p, x)[1:-1, 1:-1, 1:-1].copy()
    n_acc = n_rej = 0
    worst_u = worst_x = 0.0
    while sim.t < 2.5:
        acc, dtu, md, nan = sim.step()
        assert not nan
        n_acc += acc
        n_rej += not acc
        t = sim.t
        u = inn3(p, sim.arr("uu"))
        xi = pad3(p, sim.arr("xb"))[1:-1, 1:-1, 1:-1]
        worst_u = max(worst_u, np.abs(u - u0 * np.exp(-nu * t)).max())
        worst_x = max(worst_x, np.abs(xi - x0 - u0 * (1 - np.exp(-nu * t)) / nu).max())
    print(f"{n_acc} accepted, {n_rej} rejected; max |u - u_exact| {worst_u:.2e}, |x - x_exact| {worst_x:.2e}")
    assert n_acc > 5 and n_rej >= 1       # the first dt is far too large for maxError = 1e-10
    assert worst_u < 1e-9 and worst_x < 1e-9
    # the grid stayed a rigid translation: unit det, J = 0, U2 = |u|^2 L^3
    sim.fields()
    assert np.abs(sim.arr("J")).max() < 1e-10
    ut = u0 * np.exp(-nu * sim.t)
    assert abs(sim.umax() - np.linalg.norm(ut)) < 1e-9


def test_u2_and_umax_closed_forms():
    p = grid(inertia=True)
    X = lattice(p)
    B0 = np.zeros_like(X)
    B0[..., 2] = 1.0
    sim = fields(p, B0.reshape(-1).copy(), X.reshape(-1).copy())
    u0 = np.array([0.5, -1.25, 2.0])
    sim.arr("uu")[:] = np.tile(u0, p.ninner)
    assert abs(sim.u2() - u0 @ u0 * p.Lx * p.Ly * p.Lz) < 1e-12 * (u0 @ u0) * p.Lx * p.Ly * p.Lz
    assert abs(sim.umax() - np.linalg.norm(u0)) < 1e-15
    # a larger value at one point sets the max; a NaN elsewhere is skipped
    uu = sim.arr("uu")
    uu[3 * 17: 3 * 17 + 3] = [4.0, 0.0, 3.0]
    uu[3 * 40] = np.nan
    assert sim.umax() == 5.0
    # an interior point weighs 1, a face point 1/2, an edge 1/4, a corner 1/8
    uu[:] = 0
    w = {}
    for (i, j, k), want in (((3, 2, 4), 1.0), ((0, 2, 4), 0.5), ((0, 0, 4), 0.25), ((0, 0, 0), 0.125)):
        uu[:] = 0
        q = i + p.nx * (j + p.ny * k)
        uu[3 * q] = 2.0
        w[(i, j, k)] = sim.u2() / (4.0 * p.dx * p.dy * p.dz)
        assert abs(w[(i, j, k)] - want) < 1e-15, w
