"""Mesh and error diagnostics of the oracle (cell volume, octant wedges,
convexity, errB_1ez, errXb_XbAn, epsilon*) pinned by closed forms:
affine maps give every oriented octant wedge and the cell volume as
det(A) dx dy dz; a rotation about z by phi(r, z) has |X - R X|^2 =
2 r^2 (1 - cos phi); epsilon* of a field with J.B/B^2 linear in z is
sqrt2 times the slope."""
import numpy as np

import oracle as o
from test_oracle_kat import grid, pad3


def padded_grid(p):
    """The uniform grid on all padded points: x = O + (i - 1) h."""
    k, j, i = np.meshgrid(np.arange(p.nz + 2), np.arange(p.ny + 2), np.arange(p.nx + 2), indexing="ij")
    return np.stack([p.Ox + (i - 1) * p.dx, p.Oy + (j - 1) * p.dy, p.Oz + (k - 1) * p.dz], axis=-1)


def test_affine_maps_give_det_times_cell():
    p = grid()
    X = padded_grid(p)
    h3 = p.dx * p.dy * p.dz
    for A in (np.diag([1.0, 1.0, 1.0]),
              np.array([[1.2, 0.3, -0.1], [0.05, 0.9, 0.2], [0.1, -0.2, 1.1]]),
              np.array([[0.0, 1.0, 0.0], [1.0, 0.0, 0.0], [0.0, 0.0, 1.0]])):   # det -1: inverted
        xb = (X @ A.T + np.array([0.3, -0.2, 0.1])).reshape(-1)
        cv, co, wm, cmin, wmin = o.mesh(p, xb)
        d = np.linalg.det(A) * h3
        assert np.allclose(cv, d, rtol=1e-12) and np.allclose(wm, d, rtol=1e-12)
        assert np.all(co == np.sign(d)) and cmin == np.sign(d) and abs(wmin - d) < 1e-12 * abs(d)


def test_a_folded_point_is_flagged_locally():
    p = grid()
    X = padded_grid(p)
    i0, j0, k0 = 4, 3, 5                      # inner indices
    X[k0 + 1, j0 + 1, i0 + 1, 0] += 1.5 * p.dx   # beyond its +x neighbour
    cv, co, wm, cmin, wmin = o.mesh(p, X.reshape(-1))
    co = co.reshape(p.nz, p.ny, p.nx)
    assert cmin == -1 and wmin < 0
    assert co[k0, j0, i0] == -1 and co[k0, j0, i0 + 1] == -1
    far = np.ones_like(co, dtype=bool)
    far[max(k0 - 1, 0):k0 + 2, max(j0 - 1, 0):j0 + 2, max(i0 - 1, 0):i0 + 2] = False
    assert np.all(co[far] == 1)


def test_err_b_1ez_of_a_constant_field():
    p = grid()
    B = np.zeros((p.nz + 2, p.ny + 2, p.nx + 2, 3))
    B[...] = (0.3, -0.4, 1.0 + 1.2)
    assert abs(o.err_b_1ez(p, B.reshape(-1)) - 1.3) < 1e-13


def test_err_xb_xban_of_rotations():
    p = grid(ar=1.5, az=2.0, phi1=0.8)
    X = padded_grid(p)
    inner = X[1:-1, 1:-1, 1:-1]
    x, y, z = inner[..., 0], inner[..., 1], inner[..., 2]
    phi = 0.8 * np.exp(-(x * x + y * y) / 2.25 - z * z / 4.0)
    # the uniform grid is the analytic map rotated back by phi
    want = np.sqrt(np.sum(2 * (x * x + y * y) * (1 - np.cos(phi))) / x.size)
    assert abs(o.err_xb_xban(p, X.reshape(-1)) - want) < 1e-12 * want
    # the analytic map itself: rotation of (x, y) by -phi
    Y = X.copy()
    Y[1:-1, 1:-1, 1:-1, 0] = np.cos(phi) * x + np.sin(phi) * y
    Y[1:-1, 1:-1, 1:-1, 1] = -np.sin(phi) * x + np.cos(phi) * y
    assert o.err_xb_xban(p, Y.reshape(-1)) < 1e-15


def test_epsilon_star_of_a_linear_twist_profile():
    p = grid()
    X = padded_grid(p)
    rng = np.random.default_rng(3)
    X[1:-1, :, :, 2] += 0.2 * p.dz * rng.uniform(-1, 1, (p.nz, 1, 1))   # uneven z along the column
    B = np.zeros((p.nz + 2, p.ny + 2, p.nx + 2, 3))
    B[..., 2] = 2.0
    J = np.zeros((p.nz, p.ny, p.nx, 3))
    alpha = 0.37
    J[..., 2] = 2.0 * alpha * X[1:-1, 1:-1, 1:-1, 2]    # J.B / B^2 = J_z / 2 = alpha z
    eps = o.epsilon_star(p, X.reshape(-1), B.reshape(-1), J.reshape(-1))
    assert abs(eps - 1.414213562373095 * alpha) < 1e-12
    # a kink: the steeper segment sets the maximum
    zc = X[1:-1, p.ny // 2 + 1, p.nx // 2 + 1, 2]
    a = np.where(zc < 0, alpha * zc, 3 * alpha * zc)
    J[:, p.ny // 2, p.nx // 2, 2] = 2.0 * a
    eps = o.epsilon_star(p, X.reshape(-1), B.reshape(-1), J.reshape(-1))
    assert abs(eps - 1.414213562373095 * 3 * alpha) < 1e-12
