"""Relaxation histories of the braids: the force-freeness criterion
max|JxB|/max(|J||B|), the energy, dt and the mesh quality along a run from
the initial state, for any grid, precision and braid amplitude.

The loop is the reference's main loop (src/main.cu:259-384) run by the
device controller; every `every` attempts the run stops, the fields are
evaluated and the diagnostics of src/diagnostics.cu:12-134 are reduced:
energy B2, JMax, epsilon* (src/maths.cu:232-290), the smallest oriented
octant wedge and the convexity count (src/maths.cu:497-640).  The per-
attempt criterion comes from the controller's step log.

    python -m glemur_amd.relax_study --grid 129 --prec f32k --t-end 40 --every 500
"""
from __future__ import annotations

import json
import sys
import time

from . import PREC_F32K, PREC_F64, Context, GlemurError, braid_seeded, init_state, workloads


def braid(n, amplitude=1.0, seed=None, same_sign=False):
    """The plain braid on an n x n x (3n - 2) grid (33, 65, 129, 257 nest),
    its six amplitudes scaled by `amplitude`, or a seeded braid; same_sign
    twists all six blobs the same way (a twisted tube, not a braid)."""
    p = workloads.braid_params(n, n, 3 * n - 2, n_cycle=10 ** 8)
    if seed is not None:
        braid_seeded(p, seed, 6, 6, False)
    for k in range(p.nBlobs):
        p.blobTwist[k] *= amplitude
        if same_sign:
            p.blobTwist[k] = abs(p.blobTwist[k])
    p.redWedgeMin = p.redConvex = p.redEpsilonStar = True
    return p


def layer_profile(c, p, depths=(0, 1, 2, 4, 8, 16)):
    """max|JxB| / max(|J||B|) over the inner points at least d points away
    from every face, for each d in depths, and where max|JxB| sits (its
    inner-grid index and distance to the nearest face), from the fields of
    the current state (gm_fields)."""
    import numpy as np
    from . import FIELD_B, FIELD_J
    nx, ny, nz = p.nx, p.ny, p.nz
    B = c.download(FIELD_B).reshape(nz + 2, ny + 2, nx + 2, 3)[1:-1, 1:-1, 1:-1]
    J = c.download(FIELD_J).reshape(nz, ny, nx, 3)
    f = np.sqrt(np.sum(np.cross(J, B) ** 2, axis=-1))
    jb = np.sqrt(np.sum(J * J, axis=-1) * np.sum(B * B, axis=-1))
    k, j, i = np.unravel_index(int(np.argmax(f)), f.shape)
    dist = int(min(i, j, k, nx - 1 - i, ny - 1 - j, nz - 1 - k))
    out = {"argmax_ijk": [int(i), int(j), int(k)], "argmax_face_distance": dist}
    for d in depths:
        if 2 * d >= min(nx, ny, nz):
            break
        sl = (slice(d, nz - d), slice(d, ny - d), slice(d, nx - d))
        out[f"criterion_d{d}"] = float(f[sl].max() / jb[sl].max())
    return out


def history(p, prec, t_end, every, max_attempts, dt0=0.02, unet=None, settle=50, log=None, out=None,
            layers=False):
    """Rows of (attempts, t, dt, criterion, energy, mesh diagnostics) every
    `every` attempts until t >= t_end, dt <= dtMin or max_attempts; the
    history so far is rewritten to `out` (JSON) after every row."""
    B0, xb = init_state(p)
    rows = []
    best = (float("inf"), None)
    with Context(p, prec) as c:
        c.upload(B0, xb, 0.0, dt0)
        c.sync()
        t0 = time.perf_counter()
        warm = None
        if unet is not None:
            info = unet.warm_start(c, settle=settle)
            warm = {"used": bool(info.used), "ratio_cold": info.ratio_cold, "ratio_warm": info.ratio_warm,
                    "energy_cold": info.energy_cold, "energy_warm": info.energy_warm, "ms": info.ms}
        n_done = 0
        stop = None
        while True:
            try:
                st = c.run(max_attempts=every, n_cycle=10 ** 9, batch=100)
            except GlemurError as e:   # NaN in the error estimate: the mesh has folded
                stop = str(e)
                break
            n = int(st.n_attempt) - n_done
            n_done = int(st.n_attempt)
            for r in c.step_log(min(n, 4096)):
                if 0 <= r.ratio < best[0]:
                    best = (r.ratio, r.t)
            c.fields()
            red = c.reductions()
            row = {"attempts": n_done, "accepted": int(st.n_accept), "t": st.t, "dt": st.dt,
                   "criterion": st.ratio, "criterion_pointwise": st.ratio_pt, "energy": red.B2,
                   "JMax": red.JMax, "epsilonStar": red.epsilonStar, "wedgeMin": red.wedgeMin,
                   "convex": red.convex, "seconds": time.perf_counter() - t0}
            if layers:
                row.update(layer_profile(c, p))
            rows.append(row)
            if log:
                log(json.dumps(row))
            if out:
                with open(out, "w") as f:
                    json.dump({"grid": [p.nx, p.ny, p.nz], "min_criterion": best[0], "t_at_min": best[1],
                               "warm": warm, "rows": rows}, f)
            if st.t >= t_end or st.dt <= p.dtMin or n_done >= max_attempts or st.nan_seen:
                break
    return {"grid": [p.nx, p.ny, p.nz], "precision": "F64" if prec == PREC_F64 else "F32K",
            "min_criterion": best[0], "t_at_min": best[1], "warm": warm, "stopped_by": stop, "rows": rows}


def main(argv=None):
    import argparse
    ap = argparse.ArgumentParser(description="relaxation history of a braid")
    ap.add_argument("--grid", type=int, default=129, help="n: grid n x n x (3n - 2)")
    ap.add_argument("--prec", choices=["f32k", "f64"], default="f32k")
    ap.add_argument("--amplitude", type=float, default=1.0, help="scale of the six blob amplitudes")
    ap.add_argument("--seed", type=int, default=None, help="seeded braid instead of the plain one")
    ap.add_argument("--same-sign", action="store_true", help="all blobs twisted the same way")
    ap.add_argument("--t-end", type=float, default=40.0)
    ap.add_argument("--every", type=int, default=500)
    ap.add_argument("--max-attempts", type=int, default=200000)
    ap.add_argument("--warm", action="store_true", help="start from the learned warm start")
    ap.add_argument("--out", default=None, help="write the history as JSON here")
    ap.add_argument("--layers", action="store_true",
                    help="also the criterion away from the faces and where its maximum sits")
    a = ap.parse_args(argv)
    p = braid(a.grid, a.amplitude, a.seed, a.same_sign)
    unet = None
    if a.warm:
        from . import UNet
        from . import warmstart as ws
        unet = UNet(*ws.COARSE)
        ws.load(unet)
    h = history(p, PREC_F64 if a.prec == "f64" else PREC_F32K, a.t_end, a.every, a.max_attempts,
                unet=unet, log=lambda s: print(s, flush=True), out=a.out, layers=a.layers)
    h["amplitude"] = a.amplitude
    h["seed"] = a.seed
    h["same_sign"] = a.same_sign
    print(json.dumps({k: v for k, v in h.items() if k != "rows"}), flush=True)
    if a.out:
        with open(a.out, "w") as f:
            json.dump(h, f)
    if unet is not None:
        unet.close()
    return h


if __name__ == "__main__":
    main(sys.argv[1:])
