"""FP32-with-compensated-positions path (GM_PREC_F32K) and the batched run
driver, on the GPU.

F32K has no bitwise oracle: it is judged against the FP64 oracle by the
errors the FP32 arithmetic allows (the tolerances below are written per
field), and on the relaxed state by the energy (relative 1e-5).  The run
driver (gm_run, hipGraph replay) must reproduce step-by-step execution bit
for bit.  Full-size (129 x 129 x 385) checks use properties that do not need
the oracle: the undistorted initial grid gives B = B0 and det = 1 exactly,
and the fused FP64 step equals the per-kernel FP64 path bit for bit.
"""
import numpy as np
import pytest

from conftest import edge_mask, evolved_state, face_mask, inner_mask, small_braid

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def g(require_gpu):
    import glemur_amd
    return glemur_amd


def oracle_sim(p, B0, xb):
    import oracle as o
    return o.Sim(o.from_product(p), B0, xb)


def relerr(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


SHEAR = {"initDist": b"initShearX", "initShearA": 0.05, "initShearB": 0.02}
SHIFT = {"initDist": b"centerShift", "initShearA": 0.05}


@pytest.mark.parametrize("case", [{}, {"epsilonProf": True}, {"pressure": True, "beta": 0.05},
                                  {"pressure": True, "beta": 0.05, **SHEAR},
                                  {"pressure": True, "beta": 0.05, "tile": (7, 6, 5), **SHIFT}])
def test_f32k_fields_accuracy(g, case):
    """B, det and J of the FP32 stage kernel against the FP64 oracle on a
    distorted grid.  B and det carry FP32 rounding (max-norm relative
    2e-6); J, a difference quotient of B over the cell, to 5e-6 of max|J|;
    gradP, from second differences, to 2e-5 of max|gradP|.  The last two
    cases start from an initially distorted grid, whose x and y faces take
    their ghost values from the precomputed table (src/boundaries.cu:11-35),
    one of them on several tiles."""
    case = dict(case)
    tile = case.pop("tile", None)
    p = small_braid(**case)
    B0, xb = evolved_state(p, 4)
    sim = oracle_sim(p, B0, xb)
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
        print(f"F32K field errors: B {eB:.2e} det {ed:.2e} J {eJ:.2e}")
        assert eB < 2e-6 and ed < 2e-6 and eJ < 5e-6
        # ghost edges and corners are never written: B stays B0 (in FP32)
        e = edge_mask(p)
        want = B0.reshape(-1, 3)[e].astype(np.float32).astype(np.float64)
        assert np.array_equal(ctx.download(g.FIELD_B).reshape(-1, 3)[e], want)
        if p.pressure:
            sim.gradp()
            eG = relerr(ctx.download(g.FIELD_GRADP), sim.arr("gradP"))
            print(f"gradP {eG:.2e}")
            assert eG < 2e-5


def test_f32k_step_accuracy(g):
    """One attempted step at a dt whose error estimate is near the
    tolerance: slopes to FP32 accuracy, positions to the FP32 rounding of
    the step increment, the error estimate to 1e-3 relative plus the FP32
    rounding floor of the slopes (|x5 - x4| is a small difference of slope
    sums), and the same accept decision."""
    p = small_braid()
    B0, xb = evolved_state(p, 3)
    sim = oracle_sim(p, B0, xb)
    dt = 0.05
    sim.set_dt(dt)
    x0 = sim.arr("xb").copy()
    acc_o, _, md_o, _ = sim.step()
    assert 1e-7 < md_o < 1e-4
    with g.Context(p, g.PREC_F32K) as ctx:
        ctx.upload(B0, xb, 0.0, dt)
        s = ctx.step()
        kk_o = sim.arr("kk").reshape(6, -1)
        floor = 8 * 2.0 ** -24 * np.abs(kk_o).max()
        print(f"maxDelta F32K {s.maxDelta:.6e} FP64 {md_o:.6e} floor {floor:.1e}")
        assert abs(s.maxDelta - md_o) < 1e-3 * md_o + floor
        assert bool(s.accepted) == acc_o
        kk_g = ctx.download(g.FIELD_KK).reshape(6, -1)[:5]
        ek = [relerr(kk_g[n], kk_o[n]) for n in range(5)]
        xg = ctx.download(g.FIELD_XB)
        inc = np.max(np.abs(sim.arr("xb") - x0))
        ex = np.max(np.abs(xg - sim.arr("xb"))) / inc
        print("slopes per stage " + " ".join(f"{e:.2e}" for e in ek) + f"; positions (rel. to the increment) {ex:.2e}")
        # later stages see positions displaced by up to dt |k| ~ 5e-2 of the
        # cell, where J's difference quotients amplify FP32 rounding
        assert max(ek) < 1e-4 and ex < 1e-5


def test_f32k_fixed_dt_trajectory_tracks_fp64(g):
    """150 steps of the 33 x 33 x 97 braid at a fixed dt far inside the
    tolerance (every attempt accepted in both precisions, so both runs sit
    at the same times): the energy and the relaxation criterion differ only
    by accumulated FP32 rounding (relative 1e-6 and 1e-3), positions by
    1e-6 of the total displacement, and the F32K energy never rises.

    With an adaptive dt the two runs take their own decisions: the error
    estimate |x5 - x4| is a sum of slopes cancelling by about four orders of
    magnitude, so FP32 slopes carry it to about 1e-2 relative and attempts
    within that band of maxError may be decided differently."""
    from glemur_amd import workloads
    p = workloads.small()
    B0, xb = g.init_state(p)
    n, dt = 150, 0.01
    with g.Context(p, g.PREC_F64) as a, g.Context(p, g.PREC_F32K) as b:
        a.upload(B0, xb, 0.0, dt)
        b.upload(B0, xb, 0.0, dt)
        for it in range(n):
            t = it * dt
            a.set_time(a.time()[0], dt)
            b.set_time(b.time()[0], dt)
            sa, sb = a.step(), b.step()
            assert sa.accepted and sb.accepted and sa.maxDelta < 0.1 * p.maxError
        la, lb = a.step_log(n), b.step_log(n)
        xa, xb2 = a.download(g.FIELD_XB), b.download(g.FIELD_XB)
    assert [r.t for r in la] == [r.t for r in lb]
    ea = np.array([r.energy for r in la])
    eb = np.array([r.energy for r in lb])
    ra = np.array([r.ratio for r in la])
    rb = np.array([r.ratio for r in lb])
    de = np.max(np.abs(eb - ea) / ea)
    dr = np.max(np.abs(rb - ra) / ra)
    dx = np.max(np.abs(xb2 - xa)) / np.max(np.abs(xa - xb))
    print(f"energy {ea[0]:.8e} -> {ea[-1]:.8e}; max rel. difference energy {de:.2e} criterion {dr:.2e}; "
          f"positions {dx:.2e} of the displacement")
    assert de < 1e-6 and dr < 1e-3 and dx < 1e-6
    assert np.all(np.diff(eb) <= 0) and eb[-1] < eb[0]


@pytest.mark.parametrize("prec", ["F64", "F32K"])
def test_run_equals_step_loop(g, prec):
    """gm_run (batches enqueued without host syncs, device controller) and
    gm_step one at a time reach the same state bit for bit."""
    P = getattr(g, "PREC_" + prec)
    p = small_braid(epsilonProf=True)
    B0, xb = evolved_state(p, 1)
    with g.Context(p, P) as a, g.Context(p, P) as b:
        a.upload(B0, xb, 0.0, 0.02)
        b.upload(B0, xb, 0.0, 0.02)
        for _ in range(40):
            b.step()
        st = a.run(max_attempts=40, batch=16)
        assert st.n_attempt == 40
        assert a.time() == b.time()
        assert np.array_equal(a.download(g.FIELD_XB), b.download(g.FIELD_XB))


@pytest.mark.parametrize("prec", ["F64", "F32K"])
def test_graph_replay_equals_stream(g, prec):
    """The hipGraph of one batch, replayed several times with host syncs in
    between, gives the same trajectory as stream launches."""
    P = getattr(g, "PREC_" + prec)
    p = small_braid()
    B0, xb = evolved_state(p, 1)
    with g.Context(p, P) as a, g.Context(p, P) as b:
        a.upload(B0, xb, 0.0, 0.02)
        b.upload(B0, xb, 0.0, 0.02)
        for _ in range(4):
            sa = a.run(max_attempts=8, batch=8, use_graph=True)
            sb = b.run(max_attempts=8, batch=8, use_graph=False)
            assert (sa.n_attempt, sa.n_accept, sa.t, sa.dt) == (sb.n_attempt, sb.n_accept, sb.t, sb.dt)
            assert np.array_equal(a.download(g.FIELD_XB), b.download(g.FIELD_XB))
        la, lb = a.step_log(32), b.step_log(32)
        assert [(r.maxDelta, r.dt_used, r.accepted) for r in la] == [(r.maxDelta, r.dt_used, r.accepted) for r in lb]


def test_f32k_deterministic(g):
    """Fixed-order reductions: two identical runs agree bit for bit."""
    p = small_braid()
    B0, xb = evolved_state(p, 1)
    res = []
    for _ in range(2):
        with g.Context(p, g.PREC_F32K) as ctx:
            ctx.upload(B0, xb, 0.0, 0.02)
            st = ctx.run(max_attempts=20, batch=10)
            res.append((st.energy, st.ratio, st.dt, ctx.download(g.FIELD_XB)))
    assert res[0][:3] == res[1][:3] and np.array_equal(res[0][3], res[1][3])


def test_run_stop_ratio(g):
    """gm_run stops at the first evaluated state whose
    max|JxB|/(|J||B|) is at most stop_ratio."""
    p = small_braid()
    B0, xb = evolved_state(p, 0)
    with g.Context(p, g.PREC_F32K) as ctx:
        ctx.upload(B0, xb, 0.0, p.dt0)
        st0 = ctx.run(max_attempts=5, batch=5)
        target = st0.ratio * 0.5
        st = ctx.run(max_attempts=5000, n_cycle=100000, stop_ratio=target, batch=32)
        assert st.converged and st.ratio <= target
        n2 = int(st.n_attempt - st0.n_attempt)
        log = ctx.step_log(min(64, n2 + 1))
        assert log[-1].ratio <= target
        assert all(r.ratio > target for r in log[:-1])


# ------------------------------------------------------------ full size
@pytest.fixture(scope="module")
def medium(g):
    from glemur_amd import workloads
    p = workloads.medium()
    B0, xb = g.init_state(p)
    return p, B0, xb


@pytest.mark.parametrize("prec", ["F64", "F32K"])
def test_full_size_initial_fields_exact(g, medium, prec):
    """129 x 129 x 385, undistorted dyadic grid: every jacobian is the
    identity exactly, so B = B0 (rounded to the stage precision) and det = 1
    at every inner point."""
    p, B0, xb = medium
    with g.Context(p, getattr(g, "PREC_" + prec)) as ctx:
        ctx.upload(B0, xb, 0.0, 0.02)
        ctx.fields()
        m = inner_mask(p)
        B = ctx.download(g.FIELD_B).reshape(-1, 3)[m]
        want = B0.reshape(-1, 3)[m]
        if prec == "F32K":
            want = want.astype(np.float32).astype(np.float64)
        assert np.array_equal(B, want)
        assert np.all(ctx.download(g.FIELD_DETJAC)[m] == 1.0)


def test_full_size_fused_equals_per_kernel_fp64(g, medium):
    """At the full 129 x 129 x 385 shape the fused FP64 step (tiled 2.5D
    march) and the per-kernel FP64 path (one launch per reference kernel)
    agree bit for bit on slopes, positions and the error max."""
    p, B0, xb = medium
    dt = 0.02
    with g.Context(p, g.PREC_F64) as a:
        a.upload(B0, xb, 0.0, dt)
        a.run(max_attempts=3, batch=3)          # leave the undistorted grid
        x1 = a.download(g.FIELD_XB)
        t1, dt1 = a.time()
        s = a.step()
        kk_f = a.download(g.FIELD_KK)
        x_f = a.download(g.FIELD_XB)
    with g.Context(p, g.PREC_F64) as b:
        b.upload(B0, x1, t1, dt1)
        for n in range(6):
            b.b_jacb0()
            b.set_bbound()
            b.current()
            b.rk_kk(n, dt1)
        md = b.rk_xnewstar()
        assert md == s.maxDelta
        assert np.array_equal(b.download(g.FIELD_KK).reshape(6, -1)[:5], kk_f.reshape(6, -1)[:5])
        if s.accepted:
            b.accept()
            assert np.array_equal(b.download(g.FIELD_XB), x_f)


def test_large_pressure_braid_f32k_tracks_fp64(g):
    """The 257 x 257 x 769 braid with seeded amplitudes and pressure on
    (51 M inner points, the F32K positions near their 32-bit offset limit):
    20 steps from the initial state in FP64 and in F32K at half the dt the
    controller reaches in 12 FP64 attempts, so every attempt is accepted in
    both and the runs sit at the same times.  The energy differs by at most
    1e-6 relative, the relaxation criterion by 1e-3, the positions by 1e-6
    of the total displacement, and the F32K energy never rises."""
    from glemur_amd import workloads
    p = workloads.large()
    B0, xb = g.init_state(p)
    with g.Context(p, g.PREC_F64) as a:
        a.upload(B0, xb, 0.0, p.dt0)
        a.run(max_attempts=12, batch=12)
        dt = 0.5 * a.time()[1]
    n = 20
    print(f"fixed dt {dt:.4e}")
    with g.Context(p, g.PREC_F64) as a, g.Context(p, g.PREC_F32K) as b:
        a.upload(B0, xb, 0.0, dt)
        b.upload(B0, xb, 0.0, dt)
        for it in range(n):
            a.set_time(a.time()[0], dt)
            b.set_time(b.time()[0], dt)
            sa, sb = a.step(), b.step()
            assert sa.accepted and sb.accepted and sa.maxDelta < 0.1 * p.maxError, it
        la, lb = a.step_log(n), b.step_log(n)
        xa, xf = a.download(g.FIELD_XB), b.download(g.FIELD_XB)
    ea = np.array([r.energy for r in la])
    eb = np.array([r.energy for r in lb])
    ra = np.array([r.ratio for r in la])
    rb = np.array([r.ratio for r in lb])
    de = np.max(np.abs(eb - ea) / ea)
    dr = np.max(np.abs(rb - ra) / ra)
    dx = np.max(np.abs(xf - xa)) / np.max(np.abs(xa - xb))
    print(f"energy {ea[0]:.8e} -> {ea[-1]:.8e}; max rel. difference energy {de:.2e} criterion {dr:.2e}; "
          f"positions {dx:.2e} of the displacement")
    assert de < 1e-6 and dr < 1e-3 and dx < 1e-6
    assert np.all(np.diff(eb) <= 0) and eb[-1] < eb[0]


def test_medium_braid_f32k_tracks_fp64_to_the_relaxed_state(g, medium):
    """The 129 x 129 x 385 braid relaxed at a fixed dt of 2e-3 for 9000
    steps (every attempt accepted in both precisions) in FP64 and in F32K,
    to t = 18, where max|JxB|/max(|J||B|) reaches its floor on this grid
    (about 0.023; it rises again afterwards as the current layer on the
    axis sharpens), so the end state is the relaxed state.  At the end:
    energy within 1e-5 relative (the accuracy target of the compensated
    path), the relaxation criterion within 1e-3, B within 1e-5 and J within
    1e-3 of their maxima, positions within 1e-5 of the total displacement;
    the F32K energy never rises."""
    p, B0, xb = medium
    n, dt = 9000, 2e-3
    out = {}
    for prec in ("F64", "F32K"):
        with g.Context(p, getattr(g, "PREC_" + prec)) as c:
            c.upload(B0, xb, 0.0, dt)
            energies = []
            for it in range(n):
                c.set_time(c.time()[0], dt)
                s = c.step()
                assert s.accepted and s.maxDelta < p.maxError, (prec, it, s.maxDelta)
                if it % 100 == 0:
                    energies.append(c.step_log(1)[0].energy)
            c.fields()
            r = c.reductions()
            out[prec] = dict(x=c.download(g.FIELD_XB), B=c.download(g.FIELD_B), J=c.download(g.FIELD_J),
                             energy=r.B2, crit=r.JxBMax_JBMax, e=np.array(energies), t=c.time()[0])
    a, b = out["F64"], out["F32K"]
    m = inner_mask(p)
    de = abs(b["energy"] - a["energy"]) / a["energy"]
    dc = abs(b["crit"] - a["crit"]) / a["crit"]
    dB = relerr(b["B"].reshape(-1, 3)[m], a["B"].reshape(-1, 3)[m])
    dJ = relerr(b["J"], a["J"])
    dx = np.max(np.abs(b["x"] - a["x"])) / np.max(np.abs(a["x"] - xb))
    print(f"t = {a['t']:.3f}: energy {a['e'][0]:.8e} -> {a['energy']:.8e} (F64), criterion {a['crit']:.4e}; "
          f"F32K vs FP64: energy {de:.2e}, criterion {dc:.2e}, B {dB:.2e}, J {dJ:.2e}, positions {dx:.2e}")
    assert de < 1e-5 and dc < 1e-3 and dB < 1e-5 and dJ < 1e-3 and dx < 1e-5
    assert np.all(np.diff(b["e"]) <= 0)
    assert a["crit"] < 0.03 and abs(a["t"] - n * dt) < 1e-9


def test_f32k_download_b0_and_compensation_word(g):
    """An F32K context stores B0 as FP32 (x, y, z) triples and each position
    as its FP32 rounding plus the FP32 rounding of the remainder.  Downloads
    of an uploaded distorted state give back, bit for bit, float32(B0) for
    GM_FIELD_B0, float32(xb - float32(xb)) for GM_FIELD_XB_LO and the sum of
    the two words for GM_FIELD_XB."""
    p = small_braid()
    B0, xb = evolved_state(p, 3)
    with g.Context(p, g.PREC_F32K) as c:
        c.upload(B0, xb, 0.0, 0.01)
        b0 = c.download(g.FIELD_B0)
        lo = c.download(g.FIELD_XB_LO)
        x = c.download(g.FIELD_XB)
    hi = xb.astype(np.float32).astype(np.float64)
    want_lo = (xb - hi).astype(np.float32).astype(np.float64)
    assert np.array_equal(b0, B0.astype(np.float32).astype(np.float64))
    assert np.array_equal(lo, want_lo)
    assert np.array_equal(x, hi + want_lo)
    assert np.max(np.abs(x - xb)) <= 1e-12 * np.max(np.abs(xb))
