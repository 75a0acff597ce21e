"""Mesh and error diagnostics on the GPU (gm_reductions with the
reference's red* flags, and the per-point cell volume, convexity and
wedge minimum fields) against the oracle on the same state: the per-point
mesh measures and epsilon* use the oracle's operation order and agree bit
for bit (FP64 and F32K, whose positions are the same hi + lo doubles the
oracle receives); the two RMS errors are sums in another order (1e-12)."""
import numpy as np
import pytest

from conftest import evolved_state, small_braid

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def g(require_gpu):
    import glemur_amd
    return glemur_amd


FLAGS = dict(redConvex=True, redWedgeMin=True, redErrB_1ez=True, redErrXb_XbAn=True, redEpsilonStar=True)


def params(**kw):
    p = small_braid(**kw)
    for k, v in FLAGS.items():
        setattr(p, k, v)
    p.ar, p.az, p.phi1 = 1.3, 3.0, 0.6
    return p


@pytest.mark.parametrize("prec", ["F64", "F32K"])
def test_diagnostics_match_oracle(g, prec):
    import oracle as o
    p = params()
    B0, xb = evolved_state(p, 4)
    q = o.from_product(p)
    with g.Context(p, getattr(g, "PREC_" + prec)) as c:
        c.upload(B0, xb, 0.0, 0.01)
        c.fields()
        r = c.reductions()
        x = c.download(g.FIELD_XB)                 # hi + lo in F32K
        B = c.download(g.FIELD_B)
        J = c.download(g.FIELD_J)
        fields = [c.download(f) for f in (g.FIELD_CELLVOL, g.FIELD_CONVEXITY, g.FIELD_WEDGEMIN)]
    cv, co, wm, cmin, wmin = o.mesh(q, x)
    for got, want in zip(fields, (cv, co, wm)):
        assert np.array_equal(got, want)
    assert r.convex == cmin == 1 and r.wedgeMin == wmin > 0
    assert abs(r.errB_1ez - o.err_b_1ez(q, B)) <= 1e-12 * r.errB_1ez
    assert abs(r.errXb_XbAn - o.err_xb_xban(q, x)) <= 1e-12 * r.errXb_XbAn
    assert r.epsilonStar == o.epsilon_star(q, x, B, J) and r.epsilonStar > 0
    print(prec, "convex", r.convex, "wedgeMin", r.wedgeMin, "errB_1ez", r.errB_1ez, "errXb_XbAn", r.errXb_XbAn,
          "epsilonStar", r.epsilonStar)


def test_folded_mesh_is_reported(g):
    import oracle as o
    p = params()
    B0, xb = g.init_state(p)
    x = xb.reshape(p.nz + 2, p.ny + 2, p.nx + 2, 3).copy()
    x[8, 6, 5, 0] += 1.5 * p.dx                 # past its +x neighbour
    x = x.reshape(-1)
    with g.Context(p, g.PREC_F64) as c:
        c.upload(B0, x, 0.0, 0.01)
        c.fields()
        r = c.reductions()
        co = c.download(g.FIELD_CONVEXITY)
    _, co_o, _, cmin, wmin = o.mesh(o.from_product(p), x)
    assert r.convex == -1 == cmin and r.wedgeMin == wmin < 0
    assert np.array_equal(co, co_o) and np.sum(co < 0) >= 2
