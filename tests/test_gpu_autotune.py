"""Tile autotuner (gm_autotune): it times stage launches on the current
state and must leave that state untouched; the tiles it picks must give the
same trajectory as the default tiles (the per-point arithmetic of a stage
kernel does not depend on the tile: FP64 bit for bit, F32K bit for bit as
well, since max-reductions are order-free and the positions never pass
through a sum over workgroups)."""
import numpy as np
import pytest

from conftest import evolved_state, small_braid

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def g(require_gpu):
    import glemur_amd
    return glemur_amd


@pytest.mark.parametrize("prec", ["F64", "F32K"])
def test_autotune_keeps_state_and_trajectory(g, prec):
    p = small_braid(37, 33, 45)
    p.nCycle = 10 ** 6
    B0, xb = evolved_state(p, 2)
    pr = getattr(g, "PREC_" + prec)
    with g.Context(p, pr) as ref:
        ref.upload(B0, xb, 0.0, 0.01)
        ref.run(max_attempts=12, batch=4)
        want = (ref.download(g.FIELD_XB), ref.download(g.FIELD_XB_LO), ref.time())
    with g.Context(p, pr) as c:
        c.upload(B0, xb, 0.0, 0.01)
        before = c.download(g.FIELD_XB)
        ms = c.autotune(reps=2)
        print(prec, "autotuned stage ms", ["%.4f" % m for m in ms],
              [c.stage_tile(s) for s in range(-1, 6)], flush=True)
        assert all(0 < m < 1e3 for m in ms)
        for s in range(-1, 6):
            tx, ty, cz, nb = c.stage_tile(s)
            assert tx > 0 and ty > 0 and 0 < cz <= p.nz and nb > 0
        assert np.array_equal(c.download(g.FIELD_XB), before)
        assert c.time() == (0.0, 0.01)
        c.run(max_attempts=12, batch=4)
        got = (c.download(g.FIELD_XB), c.download(g.FIELD_XB_LO), c.time())
    assert got[2] == want[2]
    assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])


@pytest.mark.parametrize("press", [False, True])
def test_f32k_workgroup_sizes_give_the_same_trajectory(g, press):
    """F32K stage kernels at every workgroup size they are built for: on a
    61 x 57 x 40 grid the default tiles run stages 0-4 in 1024-thread and
    the final stage in 768-thread workgroups (with pressure: every stage in
    768-thread workgroups, the final stage with its per-plane sums in LDS);
    forced tiles run all stages in 256- and 512-thread workgroups.  The
    positions (hi and lo words), t and dt after 10 attempts agree bit for
    bit."""
    p = small_braid(61, 57, 40, pressure=press, beta=0.05)
    p.nCycle = 10 ** 6
    B0, xb = evolved_state(p, 1)
    res, sizes = [], []
    for tile in (None, (12, 10), (18, 16)):
        with g.Context(p, g.PREC_F32K) as c:
            c.upload(B0, xb, 0.0, 0.01)
            if tile:
                c.set_tile(tile[0], tile[1], 0)
            rects = [(c.stage_tile(s)[0] + 4) * (c.stage_tile(s)[1] + 4) for s in range(6)]
            sizes.append(rects)
            c.run(max_attempts=10, batch=5)
            res.append((c.download(g.FIELD_XB), c.download(g.FIELD_XB_LO), c.time()))
    print("rects per stage", sizes)
    if press:
        assert all(512 < r <= 768 for r in sizes[0])
    else:
        assert all(768 < r <= 1024 for r in sizes[0][:5]) and 512 < sizes[0][5] <= 768
    assert all(r <= 256 for r in sizes[1]) and all(384 < r <= 512 for r in sizes[2])
    for r in res[1:]:
        assert np.array_equal(r[0], res[0][0]) and np.array_equal(r[1], res[0][1]) and r[2] == res[0][2]
