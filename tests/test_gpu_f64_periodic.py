"""Periodic faces on the fused FP64 path.

The reference copies positions across a periodic face with the shift
-+(L + h) after every stage (setGridPeri, src/boundaries.cu:532-582) and B
without a shift (setPeri, src/boundaries.cu:450-528).  The fused FP64 stage
kernels hold a ghost column's source positions unshifted, compute its B as
its source's and shift a position where a reader sees it across a periodic
face, in the reference's operation order; every result must equal the
oracle's bit for bit: fields (B, det, J, gradP on inner and ghost-face
points), single attempts over several tile shapes, and a trajectory run on
the device's dt sequence (dt itself within 2 ulp: device and host pow).
With an initial boundary distortion the line-tied z faces' ghost values
depend on sin (addDist, src/boundaries.cu:11-35), where the device's and
glibc's may differ in the last bit: that case is held to 1e-12 relative."""
import numpy as np
import pytest

from conftest import evolved_state, face_mask, inner_mask, small_braid

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def g(require_gpu):
    import glemur_amd
    return glemur_amd


def oracle_sim(p, B0, xb):
    import oracle as o
    return o.Sim(o.from_product(p), B0, xb)


def assert_bitwise(a, b, what, rel=0.0):
    """Equal bit for bit (rel = 0), or within rel of max|b|."""
    a = np.asarray(a).reshape(-1)
    b = np.asarray(b).reshape(-1)
    if rel > 0:
        err = float(np.max(np.abs(a - b)) / np.max(np.abs(b)))
        assert err <= rel, f"{what}: {err:.3e} of max|b| (limit {rel:.0e})"
        return
    bad = np.flatnonzero(~((a == b) | (np.isnan(a) & np.isnan(b))))
    if bad.size:
        i = bad[0]
        rel = np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300))
        raise AssertionError(f"{what}: {bad.size} of {a.size} differ; first at {i}: {a[i]!r} vs {b[i]!r}; "
                             f"max rel {rel:.3e}")


P = {"pressure": True, "beta": 0.05}
CASES = {
    "x": {"xPeri": True},
    "y": {"yPeri": True},
    "z": {"zPeri": True},
    "xyz": {"xPeri": True, "yPeri": True, "zPeri": True},
    "z_pressure": {"zPeri": True, **P},
    "xy_pressure_eps": {"xPeri": True, "yPeri": True, "epsilonProf": True, **P},
    "xyz_pressure": {"xPeri": True, "yPeri": True, "zPeri": True, **P},
    # line-tied z faces with an initial boundary distortion beside periodic x
    "x_shear_pressure": {"xPeri": True, "initDist": b"initShearX", "initShearA": 0.05, "initShearB": 0.02, **P},
}


def tol(case):
    """0 (bit for bit), or the limit of the case with sin in its ghost faces."""
    return 1e-12 if "initDist" in CASES[case] else 0.0


@pytest.mark.parametrize("case", list(CASES))
def test_f64_periodic_fields_bitwise(g, case):
    """gm_fields on a distorted periodic grid: B everywhere, det on inner and
    ghost-face points, J and gradP equal the oracle's."""
    rel = tol(case)
    p = small_braid(**CASES[case])
    B0, xb = evolved_state(p, 3)
    sim = oracle_sim(p, B0, xb)
    sim.fields()
    with g.Context(p, g.PREC_F64) as ctx:
        ctx.upload(B0, xb, 0.0, 0.01)
        ctx.fields()
        m = inner_mask(p) | face_mask(p)
        assert_bitwise(ctx.download(g.FIELD_B), sim.arr("B"), "B", rel)
        assert_bitwise(ctx.download(g.FIELD_DETJAC)[m], sim.arr("detJac")[m], "detJac", rel)
        assert_bitwise(ctx.download(g.FIELD_J), sim.arr("J"), "J", rel)
        if p.pressure:
            sim.gradp()
            assert_bitwise(ctx.download(g.FIELD_GRADP), sim.arr("gradP"), "gradP", rel)


@pytest.mark.parametrize("tile", [None, (5, 3, 4), (7, 6, 23), (13, 2, 1)])
@pytest.mark.parametrize("case", ["x", "y", "z", "xyz_pressure", "x_shear_pressure"])
def test_f64_periodic_step_bitwise(g, case, tile):
    """One fused attempt (six stage launches, the controller, the ghost
    layers): stage slopes, the positions after the attempt (ghost layers
    included), error max and decision bit for bit, for tile shapes whose
    edges and z-chunk ends meet the periodic faces differently."""
    rel = tol(case)
    p = small_braid(**CASES[case])
    B0, xb = evolved_state(p, 2)
    sim = oracle_sim(p, B0, xb)
    dt = 0.011
    sim.set_dt(dt)
    acc_o, dtu_o, md_o, _ = sim.step()
    with g.Context(p, g.PREC_F64) as ctx:
        if tile:
            ctx.set_tile(*tile)
        ctx.upload(B0, xb, 0.0, dt)
        s = ctx.step()
        assert s.dt_used == dtu_o
        assert s.maxDelta == md_o if rel == 0 else abs(s.maxDelta - md_o) <= 1e-9 * md_o, (s.maxDelta, md_o)
        assert bool(s.accepted) == acc_o
        kk_g = ctx.download(g.FIELD_KK).reshape(6, -1)
        kk_o = sim.arr("kk").reshape(6, -1)
        assert_bitwise(kk_g[:5], kk_o[:5], "k_0..k_4", rel)
        assert_bitwise(ctx.download(g.FIELD_XB), sim.arr("xb"), "positions after the attempt", rel)
        assert abs(s.dt - sim.dt) <= 2 * np.spacing(sim.dt), (s.dt, sim.dt)


@pytest.mark.parametrize("case", ["xyz", "xyz_pressure", "xy_pressure_eps"])
def test_f64_periodic_trajectory_bitwise(g, case):
    """40 attempts: every attempt's error max and decision bit for bit with
    the oracle run on the device's dt sequence, the positions at every tenth
    attempt and at the end; then the fields and reductions of the result."""
    p = small_braid(**CASES[case])
    B0, xb = evolved_state(p, 0)
    sim = oracle_sim(p, B0, xb)
    with g.Context(p, g.PREC_F64) as ctx:
        ctx.upload(B0, xb, 0.0, p.dt0)
        for it in range(40):
            sim.set_dt(ctx.time()[1])
            s = ctx.step()
            acc, dtu, md, _ = sim.step()
            assert s.dt_used == dtu and s.maxDelta == md and bool(s.accepted) == acc, it
            assert abs(s.dt - sim.dt) <= 2 * np.spacing(sim.dt), it
            if it % 10 == 9:
                assert_bitwise(ctx.download(g.FIELD_XB), sim.arr("xb"), f"positions at attempt {it}")
        ctx.fields()
        r = ctx.reductions()
        sim.fields()
        assert r.JMax == sim.jmax() and r.B2 == pytest.approx(sim.b2(), rel=1e-13)


def test_f64_periodic_run_graph_and_step_loop_agree(g):
    """gm_run with stream launches, gm_run replaying a hipGraph and gm_step
    one attempt at a time reach the same periodic state bit for bit."""
    p = small_braid(xPeri=True, zPeri=True, pressure=True, beta=0.05)
    B0, xb = evolved_state(p, 1)
    with g.Context(p, g.PREC_F64) as a, g.Context(p, g.PREC_F64) as b, g.Context(p, g.PREC_F64) as c:
        for x in (a, b, c):
            x.upload(B0, xb, 0.0, 0.02)
        for _ in range(24):
            c.step()
        for _ in range(3):
            a.run(max_attempts=8, batch=8, use_graph=True)
            b.run(max_attempts=8, batch=8, use_graph=False)
        assert a.time() == b.time() == c.time()
        xa = a.download(g.FIELD_XB)
        assert np.array_equal(xa, b.download(g.FIELD_XB)) and np.array_equal(xa, c.download(g.FIELD_XB))
