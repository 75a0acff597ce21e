"""The 64-member ensemble workload on the GPU: 64 seeded braids (2-8 blobs
at seeded positions, widths and signs; seeds 1000..1063) on grids drawn
from 33x33x97, 49x49x145, 65x65x193 and 97x97x289, 62 M inner points in
all, advanced as one ensemble (one launch per stage for every member, one
controller launch) with tiles picked by the ensemble autotuner.

Each member runs the loop of src/main.cu:259-384 on its own dt and
accept/reject sequence.  Checks:
  * two members of every size equal standalone F32K contexts on the same
    per-stage tiles and z chunks bit for bit (positions, t, dt, error maxima,
    accept decisions, criterion and energy of every attempt);
  * the autotuner leaves every member's state, t and dt unchanged;
  * one member of every size, run at a fixed dt inside the tolerance,
    tracks the FP64 path within the FP32 rounding the compensated path
    allows (energy 1e-6, criterion 1e-3, positions 1e-6 of the
    displacement), and its F32K energy never rises."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N_MEMBERS = 64
ATTEMPTS = 20


@pytest.fixture(scope="module")
def g(require_gpu):
    import glemur_amd
    return glemur_amd


@pytest.fixture(scope="module")
def members():
    from glemur_amd import workloads
    ps = [workloads.ensemble_member(i) for i in range(N_MEMBERS)]
    by_size = {}
    for i, p in enumerate(ps):
        by_size.setdefault((p.nx, p.ny, p.nz), []).append(i)
    return ps, by_size


def test_member_sizes_cover_all_four_grids(members):
    ps, by_size = members
    assert sorted(by_size) == [(33, 33, 97), (49, 49, 145), (65, 65, 193), (97, 97, 289)]
    assert all(len(v) >= 2 for v in by_size.values())
    assert 2 <= min(p.nBlobs for p in ps) and max(p.nBlobs for p in ps) <= 8
    npts = sum(p.nx * p.ny * p.nz for p in ps)
    print(f"{len(ps)} members, {npts} inner points; per size: "
          + ", ".join(f"{k}: {len(v)}" for k, v in sorted(by_size.items())))
    assert 5e7 < npts < 8e7


def _log(ctx, n):
    return [(r.t, r.dt_used, r.maxDelta, r.accepted, r.nan_seen, r.ratio, r.ratio_pt, r.energy)
            for r in ctx.step_log(n)]


def test_autotuned_ensemble_members_equal_standalone_runs(g, members):
    ps, by_size = members
    sample = [i for k in sorted(by_size) for i in by_size[k][:2]]
    got, tiles = {}, {}
    with g.Batch(ps) as b:
        for i, (m, p) in enumerate(zip(b.members, ps)):
            B0, xb = g.init_state(p)
            m.upload(B0, xb, 0.0, 0.01)
        before = {i: (b.members[i].download(g.FIELD_XB), b.members[i].time()) for i in sample}
        ms = b.autotune(reps=2)
        print("ensemble autotune, stage ms:", " ".join(f"{x:.3f}" for x in ms))
        for i in sample:
            # tuning launches write slopes and candidates only
            assert b.members[i].time() == before[i][1]
            assert np.array_equal(b.members[i].download(g.FIELD_XB), before[i][0])
            tiles[i] = [b.members[i].stage_tile(s)[:3] for s in range(6)]
        stats = b.run(max_attempts=ATTEMPTS, batch=10, use_graph=True)
        assert all(s.n_attempt == ATTEMPTS and not s.nan_seen for s in stats)
        for i in sample:
            m = b.members[i]
            got[i] = (m.download(g.FIELD_XB), m.time(), _log(m, ATTEMPTS), stats[i].n_accept)
    for i in sample:
        p = ps[i]
        print(f"member {i} {p.nx}x{p.ny}x{p.nz}, {p.nBlobs} blobs: tiles {tiles[i]}, "
              f"{got[i][3]} of {ATTEMPTS} accepted, t {got[i][1][0]:.5f}")
        B0, xb = g.init_state(p)
        with g.Context(p, g.PREC_F32K) as c:
            for s in range(6):
                c.set_stage_tile(s, *tiles[i][s])
            c.upload(B0, xb, 0.0, 0.01)
            st = c.run(max_attempts=ATTEMPTS, batch=10)
            want = (c.download(g.FIELD_XB), c.time(), _log(c, ATTEMPTS), st.n_accept)
        assert got[i][3] == want[3], i
        assert got[i][1] == want[1], i
        assert got[i][2] == want[2], i
        assert np.array_equal(got[i][0], want[0]), i


def test_one_member_per_size_tracks_fp64_at_fixed_dt(g, members):
    """Fixed dt: 0.3 of what the FP64 controller reaches in 12 attempts,
    so every attempt is accepted in both precisions and the runs sit at the
    same times (a member of an ensemble equals its standalone F32K run bit
    for bit, test above)."""
    ps, by_size = members
    n = 20
    for size in sorted(by_size):
        i = by_size[size][0]
        p = ps[i]
        B0, xb = g.init_state(p)
        with g.Context(p, g.PREC_F64) as a:
            a.upload(B0, xb, 0.0, 0.01)
            a.run(max_attempts=12, batch=12)
            dt = 0.3 * a.time()[1]
        with g.Context(p, g.PREC_F64) as a, g.Context(p, g.PREC_F32K) as c:
            a.upload(B0, xb, 0.0, dt)
            c.upload(B0, xb, 0.0, dt)
            for it in range(n):
                a.set_time(a.time()[0], dt)
                c.set_time(c.time()[0], dt)
                sa, sc = a.step(), c.step()
                assert sa.accepted and sc.accepted and sa.maxDelta < 0.5 * p.maxError, (i, it)
            la, lc = a.step_log(n), c.step_log(n)
            xa, xc = a.download(g.FIELD_XB), c.download(g.FIELD_XB)
        ea = np.array([r.energy for r in la])
        ec = np.array([r.energy for r in lc])
        ra = np.array([r.ratio for r in la])
        rc = np.array([r.ratio for r in lc])
        de = np.max(np.abs(ec - ea) / ea)
        dr = np.max(np.abs(rc - ra) / ra)
        dx = np.max(np.abs(xc - xa)) / np.max(np.abs(xa - xb))
        print(f"member {i} {size}: dt {dt:.3e}, energy {ea[0]:.8e} -> {ea[-1]:.8e}; F32K vs FP64: "
              f"energy {de:.2e}, criterion {dr:.2e}, positions {dx:.2e} of the displacement")
        assert de < 1e-6 and dr < 1e-3 and dx < 1e-6, i
        assert np.all(np.diff(ec) <= 0) and ec[-1] < ec[0], i
