"""Whole-path relaxation checks on the GPU.

The reference's own test case (its src/params.in: bInit = "analytic", a
uniform field twisted by a known rotation map, 65^3 points) has a closed-form
relaxed state: the grid map xbAn and B = z^ (src/maths.cu:300-377,
src/diagnostics.cu:51-70).  Relaxation must drive errXb_XbAn and errB_1ez
down monotonically while the energy falls, in both precisions, and F32K
must track FP64.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def g(require_gpu):
    import glemur_amd
    return glemur_amd


def analytic_errors(g, ctx, p):
    """errXb_XbAn and errB_1ez as src/maths.cu:300-377 define them, on the host."""
    nx, ny, nz = p.nx, p.ny, p.nz
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    X, Y, Z = i * p.dx + p.Ox, j * p.dy + p.Oy, k * p.dz + p.Oz
    phi = np.exp(-(X ** 2 + Y ** 2) / p.ar ** 2 - Z ** 2 / p.az ** 2) * p.phi1
    xan = np.stack([np.cos(phi) * X + np.sin(phi) * Y, -np.sin(phi) * X + np.cos(phi) * Y, Z], -1)
    x = ctx.download(g.FIELD_XB).reshape(nz + 2, ny + 2, nx + 2, 3)[1:-1, 1:-1, 1:-1]
    ctx.fields()
    B = ctx.download(g.FIELD_B).reshape(nz + 2, ny + 2, nx + 2, 3)[1:-1, 1:-1, 1:-1]
    n = nx * ny * nz
    return (np.sqrt(((x - xan) ** 2).sum() / n), np.sqrt(((B - [0.0, 0.0, 1.0]) ** 2).sum() / n))


def test_reference_analytic_case_relaxes(g):
    p = g.params_read(os.path.join(HERE, "golden", "params", "reference_src_params.in"))
    assert p.get_str("bInit") == "analytic" and (p.nx, p.ny, p.nz) == (65, 65, 65)
    B0, xb = g.init_state(p)
    hist = {}
    for prec in (g.PREC_F64, g.PREC_F32K):
        rows = []
        with g.Context(p, prec) as ctx:
            ctx.upload(B0, xb, 0.0, p.dt0)
            rows.append((0.0, *analytic_errors(g, ctx, p), None))
            for _ in range(6):
                st = ctx.run(max_attempts=10 ** 6, n_cycle=100, batch=50)
                rows.append((st.t, *analytic_errors(g, ctx, p), st.energy))
        hist[prec] = rows
        print(prec, [(round(t, 3), f"{ex:.4e}", f"{eb:.4e}", e) for t, ex, eb, e in rows])
        ex = [r[1] for r in rows]
        eb = [r[2] for r in rows[1:]]
        en = [r[3] for r in rows[1:]]
        assert all(b < a for a, b in zip(ex, ex[1:])), ex
        assert all(b < a for a, b in zip(eb, eb[1:])), eb
        assert all(b < a for a, b in zip(en, en[1:])), en
        assert ex[-1] < 0.75 * ex[1]
    # the two precisions relax alike (they take their own dt sequences)
    for r64, r32 in zip(hist[g.PREC_F64][1:], hist[g.PREC_F32K][1:]):
        assert abs(r32[1] - r64[1]) < 2e-3 * r64[1] + 2e-4 * abs(r32[0] - r64[0])
        assert abs(r32[3] - r64[3]) < 1e-5 * r64[3]
