"""Ensembles (gm_batch_*): every stage of every member in one launch, one
controller launch for all members.  Each member must equal a standalone
context of its precision (F32K or FP64) run with the same tile bit for bit
(same stage kernel, same per-point arithmetic, same reduction order),
members stop on their own limits, and graph replay equals stream launches.
FP64 members also follow the oracle bit for bit, as standalone FP64
contexts do."""
import numpy as np
import pytest

from conftest import small_braid

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def g(require_gpu):
    import glemur_amd
    return glemur_amd


def members(g, pressure=False):
    """Four small braids of different sizes and blob sets (optionally all
    with the pressure term)."""
    kw = {"pressure": True, "beta": 0.05} if pressure else {}
    ps = [small_braid(19, 17, 23, **kw), small_braid(18, 16, 31, **kw), small_braid(21, 18, 17, epsilonProf=True, **kw)]
    from glemur_amd import workloads
    q = workloads.braid_params(20, 20, 41, n_cycle=10 ** 6)
    g.braid_seeded(q, 1003, 2, 8, True)
    q.Lx, q.Ly, q.Lz, q.Ox, q.Oy, q.Oz = 4.0, 4.0, 16.0, -2.0, -2.0, -8.0
    if pressure:
        q.pressure, q.beta = True, 0.05
    g.params_derive(q)
    ps.append(q)
    for p in ps:
        p.nCycle = 10 ** 6
    return ps


TILE = (18, 16, 9)   # rects of 21..22 x 20 points (> 384): 512-thread workgroups in both paths


@pytest.mark.parametrize("press", [False, True], ids=["line_tied", "pressure"])
def test_members_equal_standalone_runs(g, press):
    """Each ensemble member equals its standalone run bit for bit, without
    and with the pressure term (the pressure stage kernels).  The
    standalone F32K contexts march every odd stage down the grid, the
    members always upwards, so this also shows that the march direction
    does not change a bit of the result."""
    ps = members(g, press)
    states = [g.init_state(p) for p in ps]
    with g.Batch(ps) as b:
        for m, (B0, xb) in zip(b.members, states):
            m.set_tile(*TILE)
            m.upload(B0, xb, 0.0, 0.01)
        stats = b.run(max_attempts=25, batch=8)
        got = [(m.download(g.FIELD_XB), m.time(), [(r.maxDelta, r.dt_used, r.accepted, r.energy, r.ratio)
                                                    for r in m.step_log(25)]) for m in b.members]
    for i, (p, (B0, xb)) in enumerate(zip(ps, states)):
        with g.Context(p, g.PREC_F32K) as c:
            c.set_tile(*TILE)
            c.upload(B0, xb, 0.0, 0.01)
            st = c.run(max_attempts=25, batch=8)
            want = (c.download(g.FIELD_XB), c.time(),
                    [(r.maxDelta, r.dt_used, r.accepted, r.energy, r.ratio) for r in c.step_log(25)])
        assert stats[i].n_attempt == st.n_attempt == 25
        assert stats[i].n_accept == st.n_accept
        assert got[i][1] == want[1], i
        assert got[i][2] == want[2], i
        assert np.array_equal(got[i][0], want[0]), i


@pytest.mark.parametrize("prec", ["F32K", "F64"])
def test_members_stop_on_their_own_limits(g, prec):
    """Each member stops after its own nCycle accepted steps; finished
    members drop out of the remaining launches: the first to finish (5
    steps, while the others run to 20) ends where a standalone context of
    its precision stops after 5 steps, bit for bit."""
    pr = getattr(g, "PREC_" + prec)
    ps = members(g)
    for p, n in zip(ps, (5, 12, 20, 9)):
        p.nCycle = n
    with g.Batch(ps, pr) as b:
        for m, p in zip(b.members, ps):
            B0, xb = g.init_state(p)
            m.upload(B0, xb, 0.0, 0.01)
        stats = b.run(batch=4)
        for s, p in zip(stats, ps):
            assert s.n_accept == p.nCycle and s.stopped and not s.converged
        x0, t0 = b.members[0].download(g.FIELD_XB), b.members[0].time()
        # a further run advances every member by its own nCycle again
        stats2 = b.run(batch=4, n_cycle=0)
        for s, s2, p in zip(stats, stats2, ps):
            assert s2.n_accept == s.n_accept + p.nCycle
    with g.Context(ps[0], pr) as c:
        B0, xb = g.init_state(ps[0])
        c.upload(B0, xb, 0.0, 0.01)
        st = c.run(batch=4)
        assert st.n_accept == 5 and c.time() == t0
        assert np.array_equal(c.download(g.FIELD_XB), x0)


def test_graph_replay_equals_stream(g):
    ps = members(g)
    res = []
    for use_graph in (False, True):
        with g.Batch(ps) as b:
            for m, p in zip(b.members, ps):
                B0, xb = g.init_state(p)
                m.upload(B0, xb, 0.0, 0.01)
            for _ in range(3):
                b.run(max_attempts=6, batch=6, use_graph=use_graph)
            res.append([(m.time(), m.download(g.FIELD_XB)) for m in b.members])
    for (ta, xa), (tb, xb2) in zip(*res):
        assert ta == tb and np.array_equal(xa, xb2)


def test_default_tiles_mix_workgroup_sizes(g):
    """With their default tiles the members below need 1024-, 768- and
    512-thread workgroups for stages 0-4; each stage launch runs at the
    largest (members with smaller rects leave threads idle).  Every member
    equals its standalone run with default tiles: the same tiles, other z
    chunks (members share one chunk length per stage); chunks change the
    order of the energy sum only."""
    ps = [small_braid(52, 52, 24), small_braid(33, 30, 21), small_braid(19, 17, 23, epsilonProf=True)]
    for p in ps:
        p.nCycle = 10 ** 6
    states = [g.init_state(p) for p in ps]
    with g.Batch(ps) as b:
        tiles = [[m.stage_tile(s)[:2] for s in range(6)] for m in b.members]
        rects = [[(tx + 4) * (ty + 4) for tx, ty in t] for t in tiles]
        assert max(rects[0][:5]) > 768 and 512 < max(rects[1][:5]) <= 768 and max(rects[2]) <= 512, rects
        for m, (B0, xb) in zip(b.members, states):
            m.upload(B0, xb, 0.0, 0.01)
        stats = b.run(max_attempts=20, batch=5, use_graph=True)
        got = [(m.download(g.FIELD_XB), m.time(), m.step_log(20)) for m in b.members]
    for i, (p, (B0, xb)) in enumerate(zip(ps, states)):
        with g.Context(p, g.PREC_F32K) as c:
            assert [c.stage_tile(s)[:2] for s in range(6)] == tiles[i]
            c.upload(B0, xb, 0.0, 0.01)
            st = c.run(max_attempts=20, batch=5)
            x, t, log = c.download(g.FIELD_XB), c.time(), c.step_log(20)
        assert stats[i].n_attempt == st.n_attempt == 20 and stats[i].n_accept == st.n_accept
        assert got[i][1] == t, i
        assert np.array_equal(got[i][0], x), i
        for r, q in zip(got[i][2], log):
            assert (r.maxDelta, r.dt_used, r.accepted, r.ratio) == (q.maxDelta, q.dt_used, q.accepted, q.ratio), i
            assert abs(r.energy - q.energy) <= 1e-12 * abs(q.energy), i


# ------------------------------------------------------------------ FP64

def f64_members(**kw):
    """Three small FP64 braids of different sizes (one with the epsilon
    profile)."""
    ps = [small_braid(19, 17, 23, **kw), small_braid(18, 16, 31, **kw), small_braid(21, 18, 17, epsilonProf=True, **kw)]
    for p in ps:
        p.nCycle = 10 ** 6
    return ps


def log_rows(ctx, n):
    return [(r.maxDelta, r.dt_used, r.accepted, r.energy, r.ratio, r.jmax) for r in ctx.step_log(n)]


@pytest.mark.parametrize("press", [False, True], ids=["line_tied", "pressure"])
def test_f64_members_equal_standalone_runs(g, press):
    """FP64 members equal their standalone FP64 contexts run with the same
    tile bit for bit: positions, t, dt and every step-log row, the energy
    sum included (a member's block partials go to the slots of its
    standalone launch, so the controller sums them in the same order)."""
    ps = f64_members(**({"pressure": True, "beta": 0.05} if press else {}))
    states = [g.init_state(p) for p in ps]
    with g.Batch(ps, g.PREC_F64) as b:
        for m, (B0, xb) in zip(b.members, states):
            m.set_tile(*TILE)
            m.upload(B0, xb, 0.0, 0.01)
        stats = b.run(max_attempts=16, batch=8)
        got = [(m.download(g.FIELD_XB), m.time(), log_rows(m, 16)) for m in b.members]
    for i, (p, (B0, xb)) in enumerate(zip(ps, states)):
        with g.Context(p, g.PREC_F64) as c:
            c.set_tile(*TILE)
            c.upload(B0, xb, 0.0, 0.01)
            st = c.run(max_attempts=16, batch=8)
            want = (c.download(g.FIELD_XB), c.time(), log_rows(c, 16))
        assert stats[i].n_attempt == st.n_attempt == 16 and stats[i].n_accept == st.n_accept, i
        assert got[i][1] == want[1], i
        assert got[i][2] == want[2], i
        assert np.array_equal(got[i][0], want[0]), i


def test_f64_members_follow_the_oracle(g):
    """Every FP64 member, attempt by attempt with the default ensemble tiles:
    error maximum, dt used and decision equal the oracle's run on the
    device's dt sequence (dt itself within 2 ulp: device and host pow), and
    the positions after eight attempts equal the oracle's bit for bit."""
    import oracle as o
    ps = f64_members()
    states = [g.init_state(p) for p in ps]
    sims = [o.Sim(o.from_product(p), B0, xb) for p, (B0, xb) in zip(ps, states)]
    with g.Batch(ps, g.PREC_F64) as b:
        for m, (B0, xb) in zip(b.members, states):
            m.upload(B0, xb, 0.0, 0.02)
        for it in range(8):
            for m, sim in zip(b.members, sims):
                sim.set_dt(m.time()[1])
            b.run(max_attempts=1, batch=1)
            for i, (m, sim) in enumerate(zip(b.members, sims)):
                acc, dtu, md, _ = sim.step()
                r = m.step_log(1)[0]
                assert (r.maxDelta, r.dt_used, bool(r.accepted)) == (md, dtu, acc), (i, it)
                assert abs(m.time()[1] - sim.dt) <= 2 * np.spacing(sim.dt), (i, it)
        for i, (m, sim) in enumerate(zip(b.members, sims)):
            assert np.array_equal(m.download(g.FIELD_XB), sim.arr("xb")), i


def test_f64_autotuned_ensemble_and_graph_replay(g):
    """FP64 results do not depend on the tile: after the ensemble autotuner
    (512-point policies) the members' positions, t and decisions equal those
    of an ensemble on its default tiles, and a graph-replayed run equals
    stream launches."""
    ps = f64_members(pressure=True, beta=0.05)
    states = [g.init_state(p) for p in ps]
    res = []
    for tune, use_graph in ((False, False), (True, False), (True, True)):
        with g.Batch(ps, g.PREC_F64) as b:
            for m, (B0, xb) in zip(b.members, states):
                m.upload(B0, xb, 0.0, 0.01)
            if tune:
                ms = b.autotune(reps=2)
                assert len(ms) == 6 and all(t > 0 for t in ms)
                for m in b.members:
                    assert all((tx + 4) * (ty + 4) <= 512 for tx, ty, _, _ in (m.stage_tile(s) for s in range(6)))
            for _ in range(2):
                b.run(max_attempts=6, batch=6, use_graph=use_graph)
            res.append([(m.time(), m.download(g.FIELD_XB), [r[:3] for r in log_rows(m, 12)]) for m in b.members])
    for other in res[1:]:
        for (ta, xa, la), (tb, xb2, lb) in zip(res[0], other):
            assert ta == tb and la == lb and np.array_equal(xa, xb2)


def test_f64_ensemble_refuses_per_kernel_members(g):
    """Members that need the FP64 per-kernel path (inertia, jMethod Classic)
    are refused, as are periodic faces."""
    for kw in ({"inertia": True, "nu": 0.6}, {"xPeri": True}):
        with pytest.raises(g.GlemurError):
            g.Batch([small_braid(), small_braid(**kw)], g.PREC_F64)
    p = small_braid()
    p.set_str("jMethod", "Classic")
    with pytest.raises(g.GlemurError):
        g.Batch([small_braid(), p], g.PREC_F64)
