"""Learned warm start: training data, training, weights and the
time-to-relax comparison.

Training braids: 65 x 65 x 193 braids with seeded amplitudes
c_k = (-1)^k U(0.5, 1.5) (splitmix64 seeds 10000 + i, six blobs at the
fixed centres of the plain braid, pressure off), relaxed by the plain method
for a fixed number of attempted steps as ensembles of 64 (one stage launch
for all members), and added to the network's device-resident training set
(features of B0, displacement of the relaxed positions).  Held-out braids:
129 x 129 x 385, seeds 20000 + i, and the plain braid itself.  The network
runs on the 65 x 65 x 193 grid; on 129 x 129 x 385 its prediction is
up-sampled trilinearly (the grids nest: fine point 2i is coarse point i).
"""
from __future__ import annotations

import os
import time

import numpy as np

from . import PREC_F32K, Batch, Context, UNet, braid_seeded, init_state, workloads

COARSE = (65, 65, 193)
FINE = (129, 129, 385)
TRAIN_SEED0 = 10000
HELDOUT_SEED0 = 20000
WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "unet_braid65x65x193.npz")


def braid(seed, shape):
    """Seeded braid (amplitudes U(0.5, 1.5), alternating signs); seed None
    gives the plain braid (amplitudes +-1)."""
    p = workloads.braid_params(*shape, n_cycle=10 ** 8)
    if seed is not None:
        braid_seeded(p, seed, 6, 6, False)
    return p


def generate(u, seeds, attempts=2000, members=64, dt0=0.02, log=print):
    """Relax the training braids of `seeds` (ensembles of `members`) for
    `attempts` attempted steps each and add them to u's training set."""
    seeds = list(seeds)
    t0 = time.perf_counter()
    for k in range(0, len(seeds), members):
        ps = [braid(s, COARSE) for s in seeds[k:k + members]]
        with Batch(ps) as b:
            for m, p in zip(b.members, ps):
                B0, xb = init_state(p)
                m.upload(B0, xb, 0.0, dt0)
            b.run(max_attempts=attempts, batch=100, use_graph=True)
            for m in b.members:
                u.add_sample(m)
        log(f"relaxed {k + len(ps)}/{len(seeds)} training braids ({time.perf_counter() - t0:.1f} s)")


def save(u, path=WEIGHTS):
    """Weights as bfloat16 bits (what the convolutions use), biases as
    float32, and the output scale."""
    p = u.get_params()
    w32 = p[:-1].astype(np.float32)
    u32 = w32.view(np.uint32).astype(np.uint64)
    bf = ((u32 + 0x7FFF + ((u32 >> 16) & 1)) >> 16).astype(np.uint16)
    os.makedirs(os.path.dirname(path), exist_ok=True)
    np.savez(path, bf16=bf, scale=np.array([p[-1]]))


def load(u, path=WEIGHTS):
    with np.load(path, allow_pickle=False) as z:
        w = (z["bf16"].astype(np.uint32) << 16).view(np.float32).astype(np.float64)
        u.set_params(np.concatenate([w, z["scale"].astype(np.float64)]))


def time_to_relax(p, theta, max_attempts, unet=None, settle=50, dt0=0.02, batch=50):
    """Wall time from the uploaded initial state to the first attempted step
    whose state has max|JxB|/max(|J||B|) <= theta, cold or warm-started.
    The warm time includes the network, the up-sampling, the checks and the
    settle attempts."""
    B0, xb = init_state(p)
    with Context(p, PREC_F32K) as c:
        c.upload(B0, xb, 0.0, dt0)
        c.sync()
        t0 = time.perf_counter()
        info = None
        if unet is not None:
            info = unet.warm_start(c, settle=settle)
        st = c.run(max_attempts=max_attempts, stop_ratio=theta, batch=batch)
        c.sync()
        secs = time.perf_counter() - t0
    out = {"seconds": secs, "attempts": int(st.n_attempt), "t": st.t, "converged": bool(st.converged),
           "ratio": st.ratio, "energy": st.energy}
    if info is not None:
        out.update({"warm_used": bool(info.used), "ratio_cold": info.ratio_cold, "ratio_applied": info.ratio_applied,
                    "ratio_after_settle": info.ratio_warm, "energy_cold": info.energy_cold,
                    "energy_warm": info.energy_warm, "warm_ms": info.ms})
    return out


def main(argv=None):
    import argparse
    import json
    ap = argparse.ArgumentParser(description="train the warm-start U-Net and compare time to relax")
    ap.add_argument("--samples", type=int, default=512)
    ap.add_argument("--seed0", type=int, default=TRAIN_SEED0)
    ap.add_argument("--attempts", type=int, default=2000, help="relaxation attempts per training braid")
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--lr", type=float, default=1e-3, help="peak learning rate (cosine decay over the epochs)")
    ap.add_argument("--init", default=None, help="weights to start from (.npz)")
    ap.add_argument("--out", default=None, help="where to save the trained weights (.npz)")
    ap.add_argument("--theta", type=float, default=0.03)
    ap.add_argument("--settle", type=int, default=50)
    ap.add_argument("--eval", type=int, default=2, help="held-out seeded braids to compare (plus the plain one)")
    a = ap.parse_args(argv)
    u = UNet(*COARSE, capacity=a.samples)
    if a.init:
        load(u, a.init)
    else:
        u.init_params(1)
    if a.samples > 0:
        generate(u, range(a.seed0, a.seed0 + a.samples), attempts=a.attempts)
        for e in range(a.epochs):
            t0 = time.perf_counter()
            lr = a.lr * 0.5 * (1.0 + np.cos(np.pi * e / a.epochs))
            loss = u.train(1, batch=a.batch, lr=lr, seed=a.seed0 + e)
            print(f"epoch {e}: loss {loss:.4e} ({time.perf_counter() - t0:.1f} s)", flush=True)
        if a.out:
            save(u, a.out)
    res = []
    cases = [("plain", None)] + [(f"seed{HELDOUT_SEED0 + i}", HELDOUT_SEED0 + i) for i in range(a.eval)]
    for name, seed in cases:
        p = braid(seed, FINE)
        cold = time_to_relax(p, a.theta, 20000)
        warm = time_to_relax(p, a.theta, 20000, unet=u, settle=a.settle)
        r = {"case": name, "theta": a.theta, "cold": cold, "warm": warm,
             "gain": cold["seconds"] / warm["seconds"] if warm["converged"] else None,
             "energy_rel_diff": (warm["energy"] - cold["energy"]) / cold["energy"]}
        print(json.dumps(r), flush=True)
        res.append(r)
    u.close()
    return res


if __name__ == "__main__":
    main()
