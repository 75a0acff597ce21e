#!/usr/bin/env python3
"""Benchmark: grid-point RK-stages per second of the fused relaxation step.

Workload: the 129 x 129 x 385 braid (six twisted blobs, line-tied box
[-4,4]^2 x [-24,24], friction 1, adaptive Cash-Karp, tolerance 1e-5) on one
GPU.  A "step" is one attempted Cash-Karp step (six fused stage launches and
the device step controller) over all 6.4 M inner points, inputs resident in
HBM.  value = inner points x 6 stages x attempted steps / wall time of the
timed steps (FP32 stages with compensated positions; the FP64 path is
reported beside it).

roofline: algorithmic bytes of one stage launch (80 B per grid point in
FP32+compensation, 132 B in FP64, averaged over the six stages: x0, B0 and
the earlier slopes read once, the new slope written once) over the stage
kernel's mean duration, timed with HIP events on the library's stream,
against 8 TB/s.  traffic: HBM bytes per stage launch from the rocprofv3
FETCH_SIZE/WRITE_SIZE summary in profiles/ (calibrated against a 1 GiB
copy), reported when this run uses the tile it was measured with.

cpu_baseline: the FP64 oracle (a serial/OpenMP C restatement of the
reference step) on a bounded sample of the same workload on this host.

relaxation: attempted steps, time and force-freeness reached by the F32K
path in a bounded run from the initial state towards
max|JxB| / max(|J||B|) <= 1e-3.

ensemble: 64 seeded braids of the sizes 33x33x97 ... 97x97x289 advanced
together (one launch per stage for all members); aggregate grid-point
RK-stages per second.

large: the 257 x 257 x 769 braid (seeded amplitudes, pressure on) run
for 2000 accepted F32K steps from its initial state; aggregate grid-point
RK-stages per second over the attempts.

time_to_relax: wall time from the initial state of the 129 x 129 x 385
braid to max|JxB| / max(|J||B|) <= 0.03, cold and with the learned warm
start (U-Net inference, up-sampling, checks and 50 settle attempts
included), and the energies of the two states reached.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_stage_braid129x129x385_hbm_traffic.md")


def measured_traffic():
    """{prec: (tile tuple, bytes per launch)} from the committed counter summary."""
    try:
        lines = open(TRAFFIC_FILE).read().splitlines()
    except OSError:
        return {}
    tiles, tr = {}, {}
    for l in lines:
        f = l.split()
        if len(f) == 5 and f[0].startswith("tile_"):
            tiles[f[0][5:]] = tuple(int(x) for x in f[1:])
        elif len(f) == 2 and f[0].startswith("traffic_bytes_per_launch_"):
            tr[f[0].rsplit("_", 1)[1]] = float(f[1])
    return {k: (tiles.get(k), v) for k, v in tr.items()}


def traffic_for(traffic, prec, tile):
    t = traffic.get(prec)
    return t[1] if t and t[0] == tuple(tile) else None


def measure(g, p, B0, xb, precision, steps, warmup, dt, autotune=False):
    ctx = g.Context(p, precision)
    ctx.upload(B0, xb, 0.0, dt)
    tuned = None
    if autotune:
        tuned = {"ms": ctx.autotune(reps=5), "tiles": [ctx.stage_tile(s) for s in range(6)]}
    # warm-up attempts (not timed)
    ctx.run(max_attempts=warmup, n_cycle=10 ** 9, batch=max(1, warmup))
    # per-launch HIP-event timing of the stage kernels over `steps` attempts
    ctx.timing(True)
    ctx.run(max_attempts=steps, n_cycle=10 ** 9, batch=steps)
    ms, cnt = ctx.timing_read()
    ctx.timing(False)
    # timed region (no per-launch events): `steps` attempts enqueued back to
    # back from the same state, one host sync at each end
    ctx.upload(B0, xb, 0.0, dt)
    ctx.run(max_attempts=warmup, n_cycle=10 ** 9, batch=max(1, warmup))
    ctx.sync()
    t2 = time.perf_counter()
    st2 = ctx.run(max_attempts=steps, n_cycle=10 ** 9, batch=steps)
    t3 = time.perf_counter()
    tile = ctx.tile()
    ctx.close()
    per_stage = [ms[n] / max(1, cnt[n]) for n in range(6)]
    return {
        "seconds": t3 - t2,
        "attempts": int(st2.n_attempt - warmup),
        "accepted": int(st2.n_accept),
        "ms_stage": sum(ms[:6]) / max(1, sum(cnt[:6])),
        "ms_per_stage": per_stage,
        "ms_controller": ms[6] / max(1, cnt[6]),
        "launches": int(sum(cnt[:6])),
        "tile": tuple(tile),
        "autotune": tuned,
    }


def relaxation_probe(g, p, B0, xb, seconds):
    ctx = g.Context(p, g.PREC_F32K)
    ctx.upload(B0, xb, 0.0, p.dt0)
    ctx.sync()
    t0 = time.perf_counter()
    st = None
    while time.perf_counter() - t0 < seconds:
        st = ctx.run(max_attempts=500, n_cycle=10 ** 9, stop_ratio=1e-3, batch=100)
        if st.converged or st.dt <= p.dtMin:
            break
    sec = time.perf_counter() - t0
    ctx.close()
    return {"criterion": "max|JxB|/max(|J||B|) <= 1e-3", "precision": "F32K", "seconds": sec,
            "attempts": int(st.n_attempt), "accepted": int(st.n_accept), "t": st.t, "dt": st.dt,
            "converged": bool(st.converged), "JxBMax_JBMax": st.ratio, "JxB_JBMax_pointwise": st.ratio_pt,
            "energy": st.energy, "time_to_relax_s": sec if st.converged else None}


def ensemble(g, workloads, n, steps):
    ps = [workloads.ensemble_member(i) for i in range(n)]
    pts = sum(p.nx * p.ny * p.nz for p in ps)
    b = g.Batch(ps)
    for m, p in zip(b.members, ps):
        B0, xb = g.init_state(p)
        m.upload(B0, xb, 0.0, 0.01)
    b.run(max_attempts=3, batch=3, n_cycle=10 ** 9)             # warm-up
    t0 = time.perf_counter()
    st = b.run(max_attempts=steps, batch=steps, n_cycle=10 ** 9)
    sec = time.perf_counter() - t0
    att = sum(s.n_attempt for s in st) - 3 * n
    b.close()
    sizes = sorted({(p.nx, p.ny, p.nz) for p in ps})
    return {"members": n, "inner_points": pts, "sizes": [list(s) for s in sizes],
            "attempts_per_member": steps, "seconds": sec,
            "value": sum(p.nx * p.ny * p.nz * 6 * (s.n_attempt - 3) for p, s in zip(ps, st)) / sec,
            "unit": "grid-pt RK-stages/s", "attempts_total": int(att)}


def time_to_relax(g, theta):
    """Cold start versus the learned warm start on the same braid: wall time
    to the first attempted step with max|JxB|/max(|J||B|) <= theta."""
    from glemur_amd import warmstart as ws
    p = ws.braid(None, ws.FINE)
    cold = ws.time_to_relax(p, theta, 20000)
    with g.UNet(*ws.COARSE) as u:
        ws.load(u)
        warm = ws.time_to_relax(p, theta, 20000, unet=u, settle=50)
    return {"theta": theta, "cold_s": cold["seconds"], "cold_attempts": cold["attempts"],
            "warm_s": warm["seconds"], "warm_attempts": warm["attempts"], "warm_used": warm["warm_used"],
            "warm_start_ms": warm["warm_ms"], "gain": cold["seconds"] / warm["seconds"] if warm["converged"] else None,
            "energy_cold": cold["energy"], "energy_warm": warm["energy"],
            "energy_rel_diff": (warm["energy"] - cold["energy"]) / cold["energy"],
            "weights": "glemur_amd/data/unet_braid65x65x193.npz (2048 seeded 65x65x193 braids, 2000 attempts each)"}


def large_run(g, workloads, steps):
    """The 257 x 257 x 769 braid with seeded amplitudes and pressure on, in
    F32K: `steps` accepted adaptive steps from the initial state, enqueued
    in batches of 100 attempts with the device controller."""
    p = workloads.large()
    n_pts = p.nx * p.ny * p.nz
    t0 = time.perf_counter()
    B0, xb = g.init_state(p)
    ctx = g.Context(p, g.PREC_F32K)
    ctx.upload(B0, xb, 0.0, p.dt0)
    del B0, xb
    ctx.sync()
    t1 = time.perf_counter()
    st = ctx.run(max_attempts=10 ** 6, n_cycle=steps, batch=100)
    t2 = time.perf_counter()
    log = ctx.step_log(4096)
    first = log[0] if len(log) == st.n_attempt else None
    ctx.close()
    return {"workload": "braid 257x257x769, seeded amplitudes (seed 3), pressure on (beta = 0.1 mean|B0|^2)",
            "precision": "F32K", "inner_points": n_pts, "accepted": int(st.n_accept),
            "attempts": int(st.n_attempt), "seconds": t2 - t1, "setup_s": t1 - t0,
            "value": n_pts * 6 * st.n_attempt / (t2 - t1), "unit": "grid-pt RK-stages/s",
            "t": st.t, "dt": st.dt,
            "energy_first": first.energy if first else None, "energy": st.energy,
            "JxBMax_JBMax_first": first.ratio if first else None, "JxBMax_JBMax": st.ratio}


def cpu_baseline(p, B0, xb, budget_s=20.0):
    import oracle as o
    threads = min(16, os.cpu_count() or 1)
    o.set_threads(threads)
    sim = o.Sim(o.from_product(p), B0, xb)
    n_pts = p.nx * p.ny * p.nz
    t0 = time.perf_counter()
    n = 0
    while True:
        sim.step()
        n += 1
        if time.perf_counter() - t0 > budget_s or n >= 50:
            break
    dt = time.perf_counter() - t0
    return {
        "value": n_pts * 6 * n / dt,
        "unit": "grid-pt RK-stages/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{n} attempted FP64 steps of the 129x129x385 braid from its initial state "
                  f"(oracle, gcc -O2 -fopenmp, {threads} OpenMP threads, {dt:.1f} s)",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-f64", action="store_true")
    ap.add_argument("--relax-seconds", type=float, default=10.0)
    ap.add_argument("--ensemble", type=int, default=64, help="ensemble members (0: skip)")
    ap.add_argument("--ttr", type=float, default=0.03,
                    help="time-to-relax threshold for the cold/warm-start comparison (0: skip)")
    ap.add_argument("--large-steps", type=int, default=2000,
                    help="accepted steps of the 257x257x769 pressure-on braid (0: skip)")
    ap.add_argument("--autotune", action="store_true",
                    help="pick each stage kernel's tile by timing candidates before measuring")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run a 1 GiB device copy first (reference kernel for HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; the multi-GPU decomposition is design only "
                                   "(z slabs with 2-plane halos)", "n_gpus": args.gpus}))
        sys.exit(2)

    import glemur_amd as g
    from glemur_amd import workloads

    copy_gbs = None
    if args.copy_calibration:
        ms = g.copy_probe(1 << 30, 5)
        copy_gbs = 2 * (1 << 30) / (ms * 1e-3) / 1e9

    p = workloads.medium()
    B0, xb = g.init_state(p)
    n_pts = p.nx * p.ny * p.nz
    # start from a developed dt so the attempts sample the relaxation, not
    # the first tiny step
    dt0 = 0.02

    f32 = measure(g, p, B0, xb, g.PREC_F32K, args.steps, args.warmup, dt0, args.autotune)
    value = n_pts * 6 * f32["attempts"] / f32["seconds"]
    bpp = g.algorithmic_bytes(g.PREC_F32K)
    bytes_launch = bpp * n_pts
    achieved = bytes_launch / (f32["ms_stage"] * 1e-3) / 1e9
    stage_bytes = [(48 + 12 * n) * n_pts for n in range(5)] + [120 * n_pts]
    traffic = measured_traffic()
    traffic_f32 = traffic_for(traffic, "f32k", f32["tile"])

    out = {
        "metric": "grid-pt RK-stages/s",
        "value": value,
        "unit": "grid-pt RK-stages/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * f32["seconds"] / f32["attempts"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (stages; positions f32 hi+lo; f64 det, error norm, energy, dt)",
        "data": "synthetic (closed-form braid field, no checkpoint)",
        "config": {"workload": "braid 129x129x385 (6 blobs, line-tied, friction 1, tol 1e-5)",
                   "grid": [p.nx, p.ny, p.nz], "inner_points": n_pts,
                   "tile_tx_ty_cz_workgroups": list(f32["tile"]), "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic_f32,
                     "kernel": "k_stage32<NT,n,false> (fused Cash-Karp stage n, F32K; NT = 1024 for stages 0-4, "
                               "768 for stage 5 at the default tiles), mean of stages 0-5",
                     "algorithmic_bytes_per_launch": bytes_launch,
                     "algorithmic_bytes_per_point": bpp,
                     "mean_launch_ms": f32["ms_stage"],
                     "per_stage_ms": f32["ms_per_stage"],
                     "per_stage_frac": [b / (t * 1e-3) / 1e9 / HBM_PEAK_GBS
                                        for b, t in zip(stage_bytes, f32["ms_per_stage"])],
                     "controller_ms": f32["ms_controller"],
                     "traffic_source": os.path.relpath(TRAFFIC_FILE, ROOT) if traffic_f32 else None},
    }
    if f32["autotune"]:
        out["autotune"] = {"stage_tiles_tx_ty_cz_workgroups": f32["autotune"]["tiles"],
                           "stage_ms": f32["autotune"]["ms"]}
    if copy_gbs is not None:
        out["copy_probe_GBs"] = copy_gbs
    if not args.no_f64:
        f64 = measure(g, p, B0, xb, g.PREC_F64, max(50, args.steps // 3), args.warmup, dt0)
        b64 = g.algorithmic_bytes(g.PREC_F64) * n_pts
        a64 = b64 / (f64["ms_stage"] * 1e-3) / 1e9
        out["f64"] = {"value": n_pts * 6 * f64["attempts"] / f64["seconds"],
                      "ms_per_step": 1e3 * f64["seconds"] / f64["attempts"],
                      "achieved_GBs": a64, "frac": a64 / HBM_PEAK_GBS, "mean_launch_ms": f64["ms_stage"],
                      "per_stage_ms": f64["ms_per_stage"],
                      "traffic": traffic_for(traffic, "f64", f64["tile"])}
    if args.relax_seconds > 0:
        out["relaxation"] = relaxation_probe(g, p, B0, xb, args.relax_seconds)
    if args.ensemble > 0:
        out["ensemble"] = ensemble(g, workloads, args.ensemble, 20)
    if args.large_steps > 0:
        out["large"] = large_run(g, workloads, args.large_steps)
    if args.ttr > 0:
        out["time_to_relax"] = time_to_relax(g, args.ttr)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(p, B0, xb)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
