"""Train and evaluate the learned warm start on trajectories the solver itself generates.

Trajectories: cylinder upsets between flat dies on the 10x10x10-cell Kuhn mesh with the upset flow law, one billet
per seed with D ~ U[25, 35] mm, H/D ~ U[1.2, 1.8] and shear friction ~ U[0.2, 0.6]; 20 increments of 0.5 mm each.
Every increment after the first gives one sample per node: the features at the start of the increment and the
converged nodal velocity's deviation from the trend start 2 v_n - v_{n-1} (the network learns the correction to the
trend, in units of the targets' RMS). Seeds 0-15 train, seeds 100-103 are held out.

    python -m impressionfea_amd.train_warmstart --train --evaluate

writes impressionfea_amd/data/warmstart_mlp.npz and prints the held-out comparison (Newton iterations per increment
and ram-load difference against the previous-increment start, which is the reference's default).
"""
from __future__ import annotations

import argparse
import json
import os
import time
from dataclasses import replace

import numpy as np

from .config import upset_small
from .forging3d import Forging3DSolver
from .mlp import MlpDevice, LearnedWarmStart, init_params, node_features, target, trend_start
from .warmstart import DEFAULT_WEIGHTS

TRAIN_SEEDS = list(range(16))
TEST_SEEDS = [100, 101, 102, 103]


def billet(seed: int):
    rng = np.random.default_rng(1000 + seed)
    D = rng.uniform(25.0, 35.0)
    H = D * rng.uniform(1.2, 1.8)
    m = rng.uniform(0.2, 0.6)
    wl = upset_small(seed=seed)
    shut = H - 12.0
    cfg = replace(wl.cfg, friction=m, size=H, h=max(D / wl.nxy, H / wl.nz),
                  top=replace(wl.cfg.top, p=(shut,) + (0.0,) * 7))
    return replace(wl, cfg=cfg, diameter=D, height=H, gap0=H - shut, name=f"upset-seed{seed}")


def trajectory(wl, increments=20, warm_start=None, precision=0):
    """One billet stepped by the solver; every increment after the first yields (features at its start, converged
    velocity deviation from the trend start in units of the ram speed)."""
    s = Forging3DSolver.from_workload(wl, precision=precision, warm_start=warm_start)
    feats, targets, its, loads, starts = [], [], [], [], []
    for k in range(increments):
        f = node_features(s) if k >= 1 else None
        base = trend_start(s)
        r = s.step()
        its.append(r.iterations)
        loads.append(r.load)
        starts.append(r.stats.get("start", "previous"))
        if f is not None:
            feats.append(f)
            targets.append(target(base, s.vg, wl.cfg.ram, 1.0))
    s.close()
    return feats, targets, its, loads, starts


def build_dataset(seeds, increments=20, log=print):
    X, Y = [], []
    for sd in seeds:
        t0 = time.time()
        f, t, its, loads, _ = trajectory(billet(sd), increments)
        X += f
        Y += t
        log(f"  seed {sd}: {len(f)} samples x {f[0].shape[0]} nodes, iterations {its}, {time.time() - t0:.1f}s")
    return np.concatenate(X), np.concatenate(Y)


def train(seeds=TRAIN_SEEDS, epochs=60, batch=4096, lr=2e-3, seed=0, out=DEFAULT_WEIGHTS, log=print, increments=20):
    X, Y = build_dataset(seeds, increments=increments, log=log)
    # targets are O(1e-3) of the ram speed: train on them in units of their RMS so the BF16 network sees O(1) values
    scale = float(np.sqrt(np.mean(Y.astype(np.float64) ** 2))) or 1.0
    Y = (Y / scale).astype(np.float32)
    log(f"dataset {X.shape[0]} rows, target scale {scale:.3e} x ram")
    net = MlpDevice(init_params(seed))
    net.set_dataset(X, Y)
    rng = np.random.default_rng(seed)
    hist = []
    for ep in range(epochs):
        lr_ep = lr * (0.5 * (1 + np.cos(np.pi * ep / epochs)))
        loss = net.train_epoch(rng.permutation(X.shape[0]).astype(np.int32), batch, lr_ep)
        hist.append(loss)
        if ep % 10 == 0 or ep == epochs - 1:
            log(f"  epoch {ep}: loss {loss:.4e}")
    os.makedirs(os.path.dirname(out), exist_ok=True)
    np.savez(out, params=net.get_params(), loss=np.array(hist, np.float32), scale=np.float64(scale))
    base = float(np.mean(Y.astype(np.float64) ** 2))
    log(f"saved {out}; final loss {hist[-1]:.4e} vs the plain trend start {base:.4e}")
    return hist, base


def evaluate(seeds=TEST_SEEDS, weights=DEFAULT_WEIGHTS, increments=20, log=print):
    """Held-out comparison: Newton iterations per increment and ram load against the previous-increment start (the
    reference's default), the trend start (its ExtrapolateStart option) and the energy judge without the network."""
    ws = LearnedWarmStart.load(weights)
    rows = []
    for sd in seeds:
        wl = billet(sd)
        _, _, it0, l0, _ = trajectory(wl, increments)
        wl_trend = replace(wl, cfg=replace(wl.cfg, extrapolate_start=True))
        _, _, it1, l1, _ = trajectory(wl_trend, increments)
        _, _, it2, l2, st2 = trajectory(wl, increments, warm_start=ws)
        # the same energy judge with the network replaced by the trend start: what the judge alone gains
        _, _, it3, l3, st3 = trajectory(wl, increments, warm_start=trend_start)
        row = dict(seed=sd, iterations_previous=int(np.sum(it0[1:])), iterations_trend=int(np.sum(it1[1:])),
                   iterations_judged_without_network=int(np.sum(it3[1:])),
                   iterations_learned=int(np.sum(it2[1:])), learned_starts=int(sum(s == "learned" for s in st2)),
                   trend_starts=int(sum(s == "trend" for s in st2)),
                   max_load_rel_diff_learned=float(np.max(np.abs(np.array(l2) - l0) / np.abs(l0))),
                   max_load_rel_diff_trend=float(np.max(np.abs(np.array(l1) - l0) / np.abs(l0))))
        log(f"  held-out seed {sd}: {json.dumps(row)}")
        rows.append(row)
    return rows


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--train", action="store_true")
    ap.add_argument("--evaluate", action="store_true")
    ap.add_argument("--epochs", type=int, default=60)
    ap.add_argument("--out", default=DEFAULT_WEIGHTS)
    a = ap.parse_args()
    if a.train:
        train(epochs=a.epochs, out=a.out)
    if a.evaluate:
        rows = evaluate(weights=a.out)
        tot = {k: sum(r[k] for r in rows) for k in ("iterations_previous", "iterations_trend",
                                                     "iterations_judged_without_network", "iterations_learned",
                                                     "learned_starts", "trend_starts")}
        print(json.dumps({"held_out": rows, "totals": tot}))


if __name__ == "__main__":
    main()
