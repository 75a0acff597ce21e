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
