"""gm_rk_xnewstar_noalloc: xNewStar + findDevMax (src/rungeKutta.cu:251-353,
src/maths.cu:101-130) with its scratch in the library's static device array
instead of allocated memory.  Two calls in a row, the second on a smaller
grid, each equal to gm_rk_xnewstar on the same state bit for bit (the
error maximum and the candidate positions)."""
import numpy as np
import pytest

from conftest import evolved_state, small_braid

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def g(require_gpu):
    import glemur_amd
    return glemur_amd


def _stages(ctx, p, dt):
    """The reference's per-kernel stage sequence (src/main.cu:262-306)."""
    for n in range(6):
        ctx.b_jacb0()
        ctx.set_bbound()
        ctx.current()
        if p.pressure:
            ctx.gradp()
        ctx.rk_kk(n, dt)


def test_noalloc_xnewstar_equals_allocating_entry_point(g):
    for shape, case in (((33, 30, 41), {"pressure": True, "beta": 0.05}), ((17, 19, 23), {})):
        p = small_braid(*shape, **case)
        B0, xb = evolved_state(p, 2)
        with g.Context(p, g.PREC_F64) as c:
            c.upload(B0, xb, 0.0, 0.011)
            _stages(c, p, 0.011)
            md_na = c.rk_xnewstar_noalloc()
            x_na = c.download(g.FIELD_XB_NEW)
            md = c.rk_xnewstar()
            x = c.download(g.FIELD_XB_NEW)
        print(f"{shape}: maxDelta {md_na:.6e} (no-alloc) {md:.6e}")
        assert md_na == md and md > 0
        assert np.array_equal(x_na, x)


def test_noalloc_xnewstar_rejects_f32k_contexts(g):
    p = small_braid()
    with g.Context(p, g.PREC_F32K) as c:
        with pytest.raises(g.GlemurError):
            c.rk_xnewstar_noalloc()
