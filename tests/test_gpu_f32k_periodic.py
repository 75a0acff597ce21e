"""Periodic faces on the fused FP32 path (GM_PREC_F32K).

The reference keeps a periodic ghost layer equal to the opposite inner layer
shifted by -+(L + h) along the axis (setGridPeri, src/boundaries.cu:532-582,
after every stage) and copies B of the opposite inner layer into it (setPeri,
src/boundaries.cu:450-488).  The F32K stage kernels form those images at
every stage from the opposite side's inputs, the ghost layer's B by the
interior formula, which equals the copy in exact arithmetic.  Judged against
the FP64 oracle (fields) and the FP64 context (trajectories; its fused
kernels equal the oracle bit for bit), with the tolerances of the line-tied
F32K tests; the run driver and graph replay reproduce step-by-step execution
bit for bit."""
import numpy as np
import pytest

from conftest import evolved_state, face_mask, inner_mask, small_braid

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def g(require_gpu):
    import glemur_amd
    return glemur_amd


def relerr(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


CASES = [
    {"xPeri": True},
    {"yPeri": True},
    {"zPeri": True},
    {"xPeri": True, "yPeri": True, "zPeri": True},
    {"zPeri": True, "pressure": True, "beta": 0.05},
    {"xPeri": True, "yPeri": True, "pressure": True, "beta": 0.05, "epsilonProf": True},
    # an initial boundary distortion: the line-tied faces take the face
    # table, the periodic ones the wrapped images
    {"xPeri": True, "pressure": True, "beta": 0.05, "initDist": b"initShearX", "initShearA": 0.05,
     "initShearB": 0.02},
]


def cid(case):
    return "-".join([k for k, v in case.items() if v is True] + (["shear"] if "initDist" in case else []))


@pytest.mark.parametrize("tile", [None, (6, 5, 7)], ids=["default_tile", "tiles_6x5x7"])
@pytest.mark.parametrize("case", CASES, ids=cid)
def test_periodic_f32k_fields_accuracy(g, case, tile):
    """B, det, J (and gradP) of the F32K fields pass on a distorted periodic
    grid against the FP64 oracle: B and det to 2e-6 (max-norm relative, inner
    and face points, the periodic ghost layers included), J to 5e-6 of
    max|J|, gradP to 2e-5 of max|gradP|; with one tile and with several
    tiles and z chunks, so that the wrapped columns, rows and planes meet
    tile and chunk edges."""
    import oracle as o
    p = small_braid(**case)
    B0, xb = evolved_state(p, 4)
    sim = o.Sim(o.from_product(p), B0, xb)
    sim.fields()
    with g.Context(p, g.PREC_F32K) as ctx:
        if tile:
            ctx.set_tile(*tile)
        ctx.upload(B0, xb, 0.0, 0.01)
        ctx.fields()
        m = inner_mask(p) | face_mask(p)
        eB = relerr(ctx.download(g.FIELD_B).reshape(-1, 3)[m], sim.arr("B").reshape(-1, 3)[m])
        ed = relerr(ctx.download(g.FIELD_DETJAC)[m], sim.arr("detJac")[m])
        eJ = relerr(ctx.download(g.FIELD_J), sim.arr("J"))
        msg = f"{cid(case)} tile {tile}: B {eB:.2e} det {ed:.2e} J {eJ:.2e}"
        if p.pressure:
            sim.gradp()
            eG = relerr(ctx.download(g.FIELD_GRADP), sim.arr("gradP"))
            msg += f" gradP {eG:.2e}"
            assert eG < 2e-5, msg
        print(msg)
        assert eB < 2e-6 and ed < 2e-6 and eJ < 5e-6, msg


@pytest.mark.parametrize("case", [CASES[0], CASES[2], CASES[3], CASES[4]], ids=cid)
def test_periodic_f32k_tracks_fp64_at_fixed_dt(g, case):
    """25 steps at a fixed dt (0.3 of the dt the FP64 controller reaches in
    12 attempts, so every attempt is accepted in both precisions) in F32K and
    in FP64: energy within 1e-6 relative, the relaxation
    criterion within 1e-3, positions within 1e-6 of the displacement at every
    padded point (the periodic ghost layers kept by the F32K context after
    each step included); the F32K energy never rises."""
    p = small_braid(**case)
    B0, xb = evolved_state(p, 0)
    with g.Context(p, g.PREC_F64) as a:
        a.upload(B0, xb, 0.0, 0.01)
        a.run(max_attempts=12, batch=12)
        dt = 0.3 * a.time()[1]
    n = 25
    with g.Context(p, g.PREC_F64) as a, g.Context(p, g.PREC_F32K) as b:
        b.set_tile(6, 5, 7)
        a.upload(B0, xb, 0.0, dt)
        b.upload(B0, xb, 0.0, dt)
        for it in range(n):
            a.set_time(a.time()[0], dt)
            b.set_time(b.time()[0], dt)
            sa, sb = a.step(), b.step()
            assert sa.accepted and sb.accepted and sa.maxDelta < 0.5 * p.maxError, it
        la, lb = a.step_log(n), b.step_log(n)
        xa, xf = a.download(g.FIELD_XB), b.download(g.FIELD_XB)
    ea = np.array([r.energy for r in la])
    eb = np.array([r.energy for r in lb])
    ra = np.array([r.ratio for r in la])
    rb = np.array([r.ratio for r in lb])
    de = np.max(np.abs(eb - ea) / ea)
    dr = np.max(np.abs(rb - ra) / ra)
    dx = np.max(np.abs(xf - xa)) / np.max(np.abs(xa - xb))
    print(f"{cid(case)}: dt {dt:.3e}, energy {ea[0]:.8e} -> {ea[-1]:.8e}; F32K vs FP64: energy {de:.2e}, "
          f"criterion {dr:.2e}, positions {dx:.2e} of the displacement")
    assert de < 1e-6 and dr < 1e-3 and dx < 1e-6
    assert np.all(np.diff(eb) <= 0) and eb[-1] < eb[0]


def test_periodic_f32k_run_graph_and_step_loop_agree(g):
    """gm_run with stream launches, gm_run replaying a hipGraph and gm_step
    one attempt at a time reach the same periodic state bit for bit."""
    p = small_braid(xPeri=True, zPeri=True)
    B0, xb = evolved_state(p, 1)
    with g.Context(p, g.PREC_F32K) as a, g.Context(p, g.PREC_F32K) as b, g.Context(p, g.PREC_F32K) as c:
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


def test_periodic_f32k_tiles_within_512_points(g):
    """Periodic F32K contexts run every stage in 512-thread workgroups: their
    default tiles fit, and larger stage tiles are refused."""
    p = small_braid(33, 30, 21, xPeri=True)
    with g.Context(p, g.PREC_F32K) as ctx:
        for s in range(-1, 6):
            tx, ty, cz, nb = ctx.stage_tile(s)
            assert (tx + 4) * (ty + 4) <= 512
        with pytest.raises(g.GlemurError):
            ctx.set_stage_tile(1, 33, 20, 8)   # 888 points: a 1024-thread rect
    with g.Context(small_braid(33, 30, 21), g.PREC_F32K) as ctx:
        ctx.set_stage_tile(1, 33, 20, 8)       # fine without periodic faces
