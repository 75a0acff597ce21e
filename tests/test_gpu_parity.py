"""GPU parity against the CPU oracle, through the C ABI.

FP64: every kernel of the reference's call structure, the fused stage
kernel and whole steps must equal the oracle bit for bit (the library
follows the reference's operation order and is built without FP
contraction).  The one exception is dt after an attempt: the device
controller's pow and the host's glibc pow may differ in the last bit, so dt
is compared to 2 ulp and the oracle is then given the device's dt.
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


def assert_bitwise(a, b, what):
    a = np.asarray(a)
    b = np.asarray(b)
    bad = np.flatnonzero(~((a == b) | (np.isnan(a) & np.isnan(b))))
    if bad.size:
        i = bad[0]
        rel = np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300))
        raise AssertionError(f"{what}: {bad.size} of {a.size} differ; first at {i}: {a[i]!r} vs {b[i]!r}; "
                             f"max rel {rel:.3e}")


CASES = {
    "plain": {},
    "eps_profile": {"epsilonProf": True},
    "pressure": {"pressure": True, "beta": 0.05},
    "no_zupdate": {"zUpdate": False},
}


@pytest.mark.parametrize("case", list(CASES))
def test_per_kernel_path_bitwise(g, case):
    """B_JacB0, setBbound, JStokes, gradP, kk (six stages) and xNewStar, one
    call each, against the oracle on a distorted grid."""
    p = small_braid(**CASES[case])
    B0, xb = evolved_state(p, 3)
    sim = oracle_sim(p, B0, xb)
    dt = 0.013
    with g.Context(p, g.PREC_F64) as ctx:
        ctx.upload(B0, xb, 0.0, dt)
        ctx.b_jacb0()
        ctx.set_bbound()
        sim.fields()
        m_in, m_face = inner_mask(p), face_mask(p)
        Bg, Bo = ctx.download(g.FIELD_B).reshape(-1, 3), sim.arr("B").reshape(-1, 3)
        assert_bitwise(Bg[m_in], Bo[m_in], "B inner")
        assert_bitwise(Bg[m_face], Bo[m_face], "B ghost faces")
        dg, do = ctx.download(g.FIELD_DETJAC), sim.arr("detJac")
        assert_bitwise(dg[m_in | m_face], do[m_in | m_face], "detJac")
        ctx.current()
        assert_bitwise(ctx.download(g.FIELD_J), sim.arr("J"), "J")
        if p.pressure:
            ctx.gradp()
            sim.gradp()
            assert_bitwise(ctx.download(g.FIELD_GRADP), sim.arr("gradP"), "gradP")
        # full attempt through the per-kernel calls
        sim.set_dt(dt)
        md_o = sim.attempt()
        for n in range(6):
            if n > 0:
                ctx.b_jacb0()
                ctx.set_bbound()
                ctx.current()
                if p.pressure:
                    ctx.gradp()
            ctx.rk_kk(n, dt)
        md_g = ctx.rk_xnewstar()
        assert_bitwise(ctx.download(g.FIELD_KK), sim.arr("kk"), "kk")
        assert_bitwise(ctx.download(g.FIELD_XB), sim.arr("xb"), "x_6")
        assert_bitwise(ctx.download(g.FIELD_XB_NEW), sim.arr("xb_new"), "xb_new")
        assert md_g == md_o, (md_g, md_o)


@pytest.mark.parametrize("case", list(CASES))
def test_fused_fields_bitwise(g, case):
    """The fused stage kernel's B (inner and ghost faces), det and J for the
    current positions (gm_fields) equal the oracle's."""
    p = small_braid(**CASES[case])
    B0, xb = evolved_state(p, 3)
    sim = oracle_sim(p, B0, xb)
    sim.fields()
    with g.Context(p, g.PREC_F64) as ctx:
        ctx.upload(B0, xb, 0.0, 0.01)
        ctx.fields()
        m = inner_mask(p) | face_mask(p)
        assert_bitwise(ctx.download(g.FIELD_B).reshape(-1, 3)[m], sim.arr("B").reshape(-1, 3)[m], "B")
        assert_bitwise(ctx.download(g.FIELD_DETJAC)[m], sim.arr("detJac")[m], "detJac")
        assert_bitwise(ctx.download(g.FIELD_J), sim.arr("J"), "J")
        if p.pressure:
            sim.gradp()
            assert_bitwise(ctx.download(g.FIELD_GRADP), sim.arr("gradP"), "gradP")
        # ghost edges and corners are never written: B stays B0 there
        e = edge_mask(p)
        assert_bitwise(ctx.download(g.FIELD_B).reshape(-1, 3)[e], B0.reshape(-1, 3)[e], "B on ghost edges")


@pytest.mark.parametrize("tile", [None, (5, 3, 4), (7, 6, 23), (13, 2, 1)])
@pytest.mark.parametrize("case", list(CASES))
def test_fused_step_bitwise(g, case, tile):
    """One fused attempted step (six stage launches + controller): stage
    slopes, candidate positions, error max and accept decision bit for bit,
    for several tile shapes (tile edges and z-chunk ends move)."""
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
        assert s.maxDelta == md_o, (s.maxDelta, md_o)
        assert bool(s.accepted) == acc_o
        kk_g = ctx.download(g.FIELD_KK).reshape(6, -1)
        kk_o = sim.arr("kk").reshape(6, -1)
        assert_bitwise(kk_g[:5], kk_o[:5], "k_0..k_4")
        xg = ctx.download(g.FIELD_XB)
        # accepted: x_new; rejected: the stage-6 position, as the reference
        assert_bitwise(xg, sim.arr("xb"), "positions after the attempt")
        assert abs(s.dt - sim.dt) <= 2 * np.spacing(sim.dt), (s.dt, sim.dt)


def test_fused_trajectory_bitwise(g):
    """60 attempted steps of the small braid: every attempt's error max,
    decision and positions bit for bit with the oracle run on the device's
    dt sequence; dt itself within 2 ulp of the oracle's controller."""
    p = small_braid(epsilonProf=True)
    B0, xb = evolved_state(p, 0)
    sim = oracle_sim(p, B0, xb)
    n_rej = 0
    with g.Context(p, g.PREC_F64) as ctx:
        ctx.upload(B0, xb, 0.0, p.dt0)
        for it in range(60):
            sim.set_dt(ctx.time()[1])
            s = ctx.step()
            acc, dtu, md, _ = sim.step()
            assert s.dt_used == dtu
            assert s.maxDelta == md, (it, s.maxDelta, md)
            assert bool(s.accepted) == acc
            assert abs(s.dt - sim.dt) <= 2 * np.spacing(sim.dt)
            n_rej += not acc
            if it % 20 == 19:
                assert_bitwise(ctx.download(g.FIELD_XB), sim.arr("xb"), f"positions at attempt {it}")
        assert_bitwise(ctx.download(g.FIELD_XB), sim.arr("xb"), "final positions")
    assert n_rej >= 1, "trajectory should exercise the reject path"


def test_reductions_match_oracle(g):
    p = small_braid()
    B0, xb = evolved_state(p, 3)
    sim = oracle_sim(p, B0, xb)
    sim.fields()
    with g.Context(p, g.PREC_F64) as ctx:
        ctx.upload(B0, xb)
        ctx.fields()
        r = ctx.reductions()
    assert r.B2 == pytest.approx(sim.b2(), rel=1e-13)
    assert r.JMax == sim.jmax()
    assert r.JxB_B2Max == sim.jxb_b2_max()
    assert r.JxB_JBMax == sim.jxb_jb_max()
    assert r.JxBMax_JBMax == pytest.approx(sim.jxbmax_jbmax(), rel=1e-15)


@pytest.mark.parametrize("peri", ["xPeri", "yPeri", "zPeri"])
def test_periodic_per_kernel_bitwise(g, peri):
    """Periodic boundaries on the per-kernel path: B, detJac on the ghost
    faces, the periodic copies and one whole attempt."""
    p = small_braid(**{peri: True})
    import oracle as o
    q = o.from_product(p)
    B0, xb = o.init_state(q)
    sim = o.Sim(q, B0, xb)
    for _ in range(2):
        sim.step()
    xb = sim.arr("xb").copy()
    sim = o.Sim(q, B0, xb)
    dt = 0.01
    with g.Context(p, g.PREC_F64) as ctx:
        ctx.upload(B0, xb, 0.0, dt)
        ctx.b_jacb0()
        ctx.set_bbound()
        ctx.set_peri_B()
        ctx.current()
        sim.fields()
        m = inner_mask(p) | face_mask(p)
        assert_bitwise(ctx.download(g.FIELD_B).reshape(-1, 3)[m], sim.arr("B").reshape(-1, 3)[m], "B")
        assert_bitwise(ctx.download(g.FIELD_J), sim.arr("J"), "J")
        sim.set_dt(dt)
        md_o = sim.attempt()
        for n in range(6):
            if n > 0:
                ctx.b_jacb0()
                ctx.set_bbound()
                ctx.set_peri_B()
                ctx.current()
            ctx.rk_kk(n, dt)
        md_g = ctx.rk_xnewstar()
        assert md_g == md_o
        assert_bitwise(ctx.download(g.FIELD_XB_NEW), sim.arr("xb_new"), "xb_new")


@pytest.mark.parametrize("peri", [("xPeri",), ("yPeri",), ("zPeri",), ("xPeri", "yPeri", "zPeri")])
def test_periodic_step_and_run_bitwise(g, peri):
    """gm_step / gm_run on contexts with periodic faces (the fused kernels
    and the device controller): 30 attempts bit for bit with the oracle run
    on the device's dt sequence, dt itself within 2 ulp (device and host
    pow); gm_fields + gm_reductions on the result equal the oracle's; gm_run
    honours its attempt limit."""
    p = small_braid(**{k: True for k in peri})
    B0, xb = evolved_state(p, 0)
    sim = oracle_sim(p, B0, xb)
    with g.Context(p, g.PREC_F64) as ctx:
        ctx.upload(B0, xb, 0.0, p.dt0)
        n_rej = 0
        for it in range(30):
            sim.set_dt(ctx.time()[1])
            s = ctx.step()
            acc, dtu, md, _ = sim.step()
            assert s.dt_used == dtu and s.maxDelta == md and bool(s.accepted) == acc, it
            assert abs(s.dt - sim.dt) <= 2 * np.spacing(sim.dt), it
            n_rej += not acc
        assert_bitwise(ctx.download(g.FIELD_XB), sim.arr("xb"), "positions")
        ctx.fields()
        r = ctx.reductions()
        sim.fields()
        assert r.JMax == sim.jmax() and r.B2 == pytest.approx(sim.b2(), rel=1e-13)
        st = ctx.run(max_attempts=7, batch=4)
        assert st.n_attempt == 37 and st.stopped and st.ratio > 0
        log = ctx.step_log(7)
        assert len(log) == 7 and all(x.dt_used > 0 and x.ratio > 0 for x in log)


@pytest.mark.parametrize("method", ["Classic", "Stokes4th", "jxbAver"])
def test_jmethod_variants_fields_and_steps_bitwise(g, method):
    """jMethod = Classic (src/current.cu:10-219), jMethod = Stokes4th
    (src/current.cu:413-735) and jxbAver with weight 0.3
    (src/rungeKutta.cu:357-443) on the FP64 per-kernel path: J on a
    distorted grid and 20 attempted steps (gm_step) bit for bit with the
    oracle; F32K contexts refuse them."""
    p = small_braid()
    if method == "jxbAver":
        p.jxbAver, p.jxbAverWeight = True, 0.3
    else:
        p.set_str("jMethod", method)
    B0, xb = evolved_state(p, 2)
    sim = oracle_sim(p, B0, xb)
    sim.fields()
    with g.Context(p, g.PREC_F64) as ctx:
        ctx.upload(B0, xb, 0.0, 0.01)
        ctx.fields()
        assert_bitwise(ctx.download(g.FIELD_J), sim.arr("J"), f"J ({method})")
        sim.set_dt(0.01)
        for it in range(20):
            s = ctx.step()
            acc, dtu, md, _ = sim.step()
            assert s.maxDelta == md and bool(s.accepted) == acc and s.dt == sim.dt, it
        assert_bitwise(ctx.download(g.FIELD_XB), sim.arr("xb"), f"positions ({method})")
    with pytest.raises(g.GlemurError):
        g.Context(p, g.PREC_F32K)


def test_braid_33x33x97_200_steps_tracks_oracle(g):
    """The plain FP64 workload on its full size: the 33 x 33 x 97 braid of
    line-tied faces, pressure off, friction 1, adaptive Cash-Karp at
    tolerance 1e-5, 200 accepted steps (src/main.cu:259-384).  gm_run drives
    the device controller without host round trips; the oracle runs the
    same loop on its own dt sequence.  dt agrees to 2 ulp per step, and the
    runs separate only by rounding that the controller feeds back: the error
    estimate is a difference of slopes cancelling by about four orders of
    magnitude, and dt takes 0.2 of its relative error.  Bounds: every
    attempt's decision equal; dt and the relaxation criterion within 1e-10
    relative; energy within 1e-12; error estimate within 1e-7 of the
    tolerance; final positions within 1e-10 of the box size."""
    from glemur_amd import workloads
    p = workloads.small()
    assert (p.nx, p.ny, p.nz, p.nCycle, p.maxError) == (33, 33, 97, 200, 1e-5)
    B0, xb = g.init_state(p)
    sim = oracle_sim(p, B0, xb)
    rows = []
    while sum(r[0] for r in rows) < 200:
        sim.fields()
        e, crit = sim.b2(), sim.jxbmax_jbmax()
        acc, dtu, md, nan = sim.step()
        assert not nan
        rows.append((acc, dtu, md, e, crit))
    with g.Context(p, g.PREC_F64) as ctx:
        ctx.upload(B0, xb, 0.0, p.dt0)
        st = ctx.run(max_attempts=10 ** 6, n_cycle=200, batch=25)
        log = ctx.step_log(len(rows) + 8)
        xg = ctx.download(g.FIELD_XB)
    assert len(log) == len(rows), (len(log), len(rows))
    o = np.array([r[1:] for r in rows])
    d = np.array([(r.dt_used, r.maxDelta, r.energy, r.ratio) for r in log])
    assert [bool(r.accepted) for r in log] == [r[0] for r in rows]
    rel = np.abs(d - o) / np.abs(o)
    ddt, dmd, den, dcr = rel[:, 0].max(), np.abs(d[:, 1] - o[:, 1]).max(), rel[:, 2].max(), rel[:, 3].max()
    dx = np.max(np.abs(xg - sim.arr("xb")))
    n_rej = sum(not r[0] for r in rows)
    print(f"{len(rows)} attempts ({n_rej} rejected), t = {st.t:.6f}, energy {rows[0][3]:.10e} -> "
          f"{log[-1].energy:.10e}, criterion {rows[0][4]:.4e} -> {log[-1].ratio:.4e}; max differences: "
          f"dt {ddt:.2e} rel, error estimate {dmd:.2e} abs, energy {den:.2e} rel, criterion {dcr:.2e} rel, "
          f"positions {dx:.2e} abs")
    assert ddt <= 1e-10 and dcr <= 1e-10 and den <= 1e-12
    assert dmd <= 1e-7 * p.maxError
    assert dx <= 1e-10 * p.Lz
    assert log[-1].energy < log[0].energy


@pytest.mark.parametrize("case", [{}, {"pressure": True, "beta": 0.05}, {"zPeri": True}])
def test_inertia_steps_bitwise(g, case):
    """inertia = t (src/rungeKutta.cu:112-353, inertia branches) on the FP64
    per-kernel path, from a seeded nonzero velocity: 25 attempted steps
    (gm_step) with every decision, error estimate and dt bit for bit; the
    36N stage slopes, positions and velocity after each of the first three
    attempts and at the end equal the oracle's; U2 and UMax of the result
    match under the reference's redU2/redUMax flags (U2 summed in another
    order: 1e-13); gm_run continues the same trajectory; F32K contexts
    refuse inertia."""
    p = small_braid(inertia=True, nu=0.6, redU2=True, redUMax=True, **case)
    B0, xb = evolved_state(p, 0)
    rng = np.random.default_rng(12)
    u0 = 0.05 * rng.standard_normal(3 * p.ninner)
    sim = oracle_sim(p, B0, xb)
    sim.arr("uu")[:] = u0
    n_rej = 0
    with g.Context(p, g.PREC_F64) as ctx:
        ctx.upload(B0, xb, 0.0, p.dt0)
        assert np.all(ctx.download(g.FIELD_UU) == 0)      # upload starts at rest
        ctx.upload_velocity(u0)
        assert_bitwise(ctx.download(g.FIELD_UU), u0, "uploaded velocity")
        sim.set_dt(p.dt0)
        for it in range(25):
            s = ctx.step()
            acc, dtu, md, _ = sim.step()
            assert s.dt_used == dtu and s.maxDelta == md and bool(s.accepted) == acc and s.dt == sim.dt, it
            n_rej += not acc
            if it < 3 or it == 24:
                assert_bitwise(ctx.download(g.FIELD_KK), sim.arr("kk"), f"stage slopes at attempt {it}")
                assert_bitwise(ctx.download(g.FIELD_XB), sim.arr("xb"), f"positions at attempt {it}")
                assert_bitwise(ctx.download(g.FIELD_UU), sim.arr("uu"), f"velocity at attempt {it}")
        ctx.fields()
        r = ctx.reductions()
        sim.fields()
        assert r.UMax == sim.umax() and r.U2 == pytest.approx(sim.u2(), rel=1e-13)
        assert r.U2 > 0 and r.B2 == pytest.approx(sim.b2(), rel=1e-13)
        st = ctx.run(max_attempts=6, batch=3)
        for _ in range(6):
            sim.step()
        assert st.n_attempt == 31
        assert_bitwise(ctx.download(g.FIELD_UU), sim.arr("uu"), "velocity after gm_run")
    assert n_rej >= 1, "the trajectory should exercise the reject path"
    with pytest.raises(g.GlemurError):
        g.Context(p, g.PREC_F32K)


@pytest.mark.parametrize("shape", [(2, 2, 2), (2, 3, 9), (7, 2, 3), (3, 5, 2)])
def test_tiny_and_thin_grids(g, shape):
    """Grids at the minimum of two inner points per axis and thin slabs,
    where every point touches a line-tied face: 8 attempted steps of the
    fused FP64 step (the oracle driven with the device's dt) and of the
    per-kernel sequence on the same dt sequence, bit for bit with the oracle
    (decisions, error estimates, positions); 8 F32K steps at a fixed dt far
    inside the tolerance stay within 1e-5 of the displacement (plus 1e-8 of
    the coordinates) of the oracle's positions.  (Adaptive F32K attempts within about 10% of
    maxError may be decided differently: the error estimate cancels the
    slopes by orders of magnitude and carries their FP32 rounding.)"""
    p = small_braid(*shape)
    B0, xb = evolved_state(p, 0)
    sim = oracle_sim(p, B0, xb)
    seq, n_rej = [], 0
    with g.Context(p, g.PREC_F64) as a:
        a.upload(B0, xb, 0.0, 0.01)
        for it in range(8):
            t, dt = a.time()
            seq.append((t, dt))
            sim.set_dt(dt)
            sa = a.step()
            acc, dtu, md, nan = sim.step()
            assert not nan and bool(sa.accepted) == acc and sa.maxDelta == md and sa.dt_used == dtu, (shape, it)
            n_rej += not acc
        assert_bitwise(a.download(g.FIELD_XB), sim.arr("xb"), f"fused positions {shape}")
    # F32K: 8 steps at dt = 0.01 (far inside the tolerance, all accepted in
    # both precisions) against the oracle at the same dt
    sim3 = oracle_sim(p, B0, xb)
    with g.Context(p, g.PREC_F32K) as b:
        b.upload(B0, xb, 0.0, 0.01)
        for it in range(8):
            b.set_time(b.time()[0], 0.01)
            sim3.set_dt(0.01)
            sb = b.step()
            acc, _, md, _ = sim3.step()
            assert sb.accepted and acc and md < 0.1 * p.maxError, (shape, it)
        xf = b.download(g.FIELD_XB)
    x3 = sim3.arr("xb")
    # relative to the displacement, with a floor of 1e-8 of the coordinates
    # (50 times below one FP32 ulp of them):
    # on the 2 x 2 x 2 braid the grid moves by 3e-8 in these steps, and the
    # FP32 Lorentz force there is a small difference of O(1) terms
    dev, disp = np.max(np.abs(xf - x3)), np.max(np.abs(x3 - xb))
    print(shape, f"F32K positions {dev:.2e} (displacement {disp:.2e})")
    assert dev <= 1e-5 * disp + 1e-8 * np.max(np.abs(x3))
    sim2 = oracle_sim(p, B0, xb)
    with g.Context(p, g.PREC_F64) as c:
        c.upload(B0, xb, 0.0, 0.01)
        for it, (t, dt) in enumerate(seq):
            c.set_time(t, dt)
            sim2.set_dt(dt)
            for n in range(6):
                c.b_jacb0()
                c.set_bbound()
                c.current()
                c.rk_kk(n, dt)
            m = c.rk_xnewstar()
            acc, _, md, _ = sim2.step()
            assert m == md and (m < p.maxError) == acc, (shape, it)
            if acc:
                c.accept()
        assert_bitwise(c.download(g.FIELD_XB), sim2.arr("xb"), f"per-kernel positions {shape}")
    print(shape, f"{n_rej} of 8 attempts rejected")
