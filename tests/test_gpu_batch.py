"""Ensembles (gm_batch_*): every stage of every member in one launch, one
controller launch for all members.  Each member must equal a standalone
FP32 context run with the same tile bit for bit (same stage kernel, same
per-point arithmetic, same reduction order), members stop on their own
limits, and graph replay equals stream launches."""
import numpy as np
import pytest

from conftest import small_braid

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def g(require_gpu):
    import glemur_amd
    return glemur_amd


def members(g):
    """Four small braids of different sizes and blob sets."""
    ps = [small_braid(19, 17, 23), small_braid(18, 16, 31), small_braid(21, 18, 17, epsilonProf=True)]
    from glemur_amd import workloads
    q = workloads.braid_params(20, 20, 41, n_cycle=10 ** 6)
    g.braid_seeded(q, 1003, 2, 8, True)
    q.Lx, q.Ly, q.Lz, q.Ox, q.Oy, q.Oz = 4.0, 4.0, 16.0, -2.0, -2.0, -8.0
    g.params_derive(q)
    ps.append(q)
    for p in ps:
        p.nCycle = 10 ** 6
    return ps


TILE = (18, 16, 9)   # rects of 21..22 x 20 points (> 384): 512-thread workgroups in both paths


def test_members_equal_standalone_runs(g):
    """Each ensemble member equals its standalone run bit for bit.  The
    standalone F32K contexts march every odd stage down the grid, the
    members always upwards, so this also shows that the march direction
    does not change a bit of the result."""
    ps = members(g)
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


def test_members_stop_on_their_own_limits(g):
    """Each member stops after its own nCycle accepted steps; finished
    members drop out of the remaining launches (their state stays put)."""
    ps = members(g)
    for p, n in zip(ps, (5, 12, 20, 9)):
        p.nCycle = n
    with g.Batch(ps) as b:
        for m, p in zip(b.members, ps):
            B0, xb = g.init_state(p)
            m.upload(B0, xb, 0.0, 0.01)
        stats = b.run(batch=4)
        for s, p in zip(stats, ps):
            assert s.n_accept == p.nCycle and s.stopped and not s.converged
        # a further run advances every member by its own nCycle again
        stats2 = b.run(batch=4, n_cycle=0)
        for s, s2, p in zip(stats, stats2, ps):
            assert s2.n_accept == s.n_accept + p.nCycle


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
