"""Explorer, selector training and evaluation.

    python -m vldb_voila_amd.explore --train 4000 --heldout 500 --out DIR

1. Explorer: draws pipelines from the batched workload's distribution
   (probe / build rows and group counts log-uniform, selectivity uniform,
   key skew theta in {0, 0.5, 1, 1.5}; sizes shrunk to probe rows in
   [1e4, 5e5] and build rows in [1e3, 1e5] so 4,500 x 48 timings fit a few
   minutes) and times all 48 kernel variants plus the built-in default of
   each on the GPU.  The reference's explorer does the same over CPU blends
   and stores the timings in SQLite (explorer.cpp:33-86, compiler.cpp:124-255).
2. Trains the selector MLP on the training pipelines with the HIP kernels.
3. Evaluates on the held-out pipelines: the selector's choice (exhaustive
   fallback below 0.5 confidence) against the exhaustive best, the built-in
   default heuristic, and the best fixed blend: the one variant (applicable
   to every pipeline) with the lowest total over the training pipelines,
   used for all of them -- VOILA's default blend, one flavor for every
   pipeline (QueryConfig::default_blend, runtime_framework.hpp:78), as the
   explorer's base mode picks it (explorer.cpp:112-134).

Writes timings.npz, model.npz and report.json into DIR, printing progress.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

from . import datagen, runlog, runtime, selector, workloads

PROBE_RANGE = (1e4, 5e5)
BUILD_RANGE = (1e3, 1e5)
GROUP_RANGE = (1, 1e6)


def explore(count, seed, reps=2, log_every=100, out=None, log=None):
    params = datagen.batch_params(count, seed=seed, probe_range=PROBE_RANGE, build_range=BUILD_RANGE,
                                  group_range=GROUP_RANGE)
    cfgs = selector.variant_configs() + [runtime.make_config()]   # 48 variants + the default
    X = np.zeros((count, selector.N_FEATURES), dtype=np.float32)
    T = np.zeros((count, len(cfgs)), dtype=np.float32)
    stream = runtime.Stream()
    t0 = time.time()
    for i, prm in enumerate(params):
        prm = dict(prm)
        prm["groups"] = min(prm["groups"], prm["probe_rows"])
        b, p = datagen.gen_batch_pipeline(prm, seed=seed, index=i)
        db, dp = runtime.upload(b), runtime.upload(p)
        d = workloads.batch_pipeline(prm, b, p, runtime.addrs(db), runtime.addrs(dp)).desc()
        X[i] = selector.features(prm)
        T[i] = selector.time_variants(d, reps=reps, stream=stream, configs=cfgs)
        if log is not None:
            # one row per variant in the reference's run log: the pipeline is
            # the query, the variant its blend (the default last, with mode
            # explore_default)
            q = f"pipeline_s{seed}_{i}"
            for v, c in enumerate(cfgs):
                t = float(T[i, v])
                runlog.record(log, query=q, blend=runlog.blend_string(c), time_ms=t if t > 0 else None,
                              tuples=float(prm["probe_rows"] + prm["build_rows"]),
                              result="OK" if t > 0 else "CRASH",
                              mode="explore" if v < len(cfgs) - 1 else "explore_default")
            log.commit()
        del db, dp
        if (i + 1) % log_every == 0:
            print(f"explore seed={seed}: {i + 1}/{count} pipelines, {time.time() - t0:.0f}s", flush=True)
            if out:
                np.savez(out, X=X[:i + 1], T=T[:i + 1])
    return X, T, params


def fixed_blend(T_train, T_eval):
    """Index of the variant that applies to every pipeline of both sets and
    has the lowest total time over the training pipelines."""
    vt, ve = T_train[:, :selector.N_VARIANTS], T_eval[:, :selector.N_VARIANTS]
    ok = (vt > 0).all(0) & (ve > 0).all(0)
    tot = np.where(vt > 0, vt, np.inf).sum(0)
    return int(np.argmin(np.where(ok, tot, np.inf)))


def evaluate(net, mu, sd, X, T, params, fixed=None):
    """Selector decision per held-out pipeline, priced with the measured
    timings: chosen variant (or the exhaustive best when it falls back).
    fixed: the fixed blend's variant index (see fixed_blend)."""
    var = T[:, :selector.N_VARIANTS]
    default = T[:, selector.N_VARIANTS]
    mask = (var > 0).astype(np.uint8)
    best = np.where(var > 0, var, np.inf).min(1)
    p = net.predict((X - mu) / sd, mask)
    pick = p.argmax(1)
    conf = p[np.arange(len(p)), pick]
    fallback = (conf < selector.CONFIDENCE) | (mask[np.arange(len(p)), pick] == 0)
    chosen = np.where(fallback, best, var[np.arange(len(p)), pick])
    theta = np.array([q["theta"] for q in params])
    skew = theta >= 1.0
    extra = {}
    if fixed is not None:
        f = var[:, fixed]
        extra = {"fixed_blend": {"variant": fixed, "total_ms": float(f.sum()),
                                 "over_selector_total": float(f.sum() / chosen.sum()),
                                 "over_selector_skewed": float(f[skew].sum() / chosen[skew].sum()) if skew.any() else None,
                                 "over_best_total": float(f.sum() / best.sum())}}
    return dict(extra, **{
        "heldout": int(len(X)),
        "fallback_fraction": float(fallback.mean()),
        "chosen_over_best_mean": float(np.mean(chosen / best)),
        "within_10pct_of_best": float(np.mean(chosen <= 1.1 * best)),
        "total_ms": {"selector": float(chosen.sum()), "exhaustive_best": float(best.sum()),
                     "default": float(default.sum())},
        "selector_vs_best_total": float(chosen.sum() / best.sum()),
        "default_over_selector_total": float(default.sum() / chosen.sum()),
        "default_over_selector_skewed": float(default[skew].sum() / chosen[skew].sum()) if skew.any() else None,
        "default_over_best_skewed": float(default[skew].sum() / best[skew].sum()) if skew.any() else None,
        "top1_accuracy": float(np.mean(pick == np.where(var > 0, var, np.inf).argmin(1))),
    })


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--train", type=int, default=4000)
    ap.add_argument("--heldout", type=int, default=500)
    ap.add_argument("--epochs", type=int, default=300)
    ap.add_argument("--out", default="selector_out")
    ap.add_argument("--reps", type=int, default=2)
    ap.add_argument("--timings", default=None, help="reuse an existing timings.npz")
    ap.add_argument("--runlog", default=None, help="also record every timing in this SQLite run log "
                                                    "(the reference's runs table)")
    a = ap.parse_args(argv)
    os.makedirs(a.out, exist_ok=True)
    runtime.init(0)
    if a.timings:
        z = np.load(a.timings)
        Xtr, Ttr, Xho, Tho = z["Xtr"], z["Ttr"], z["Xho"], z["Tho"]
        pho = datagen.batch_params(len(Xho), seed=4242, probe_range=PROBE_RANGE, build_range=BUILD_RANGE,
                                   group_range=GROUP_RANGE)
        ptr = datagen.batch_params(len(Xtr), seed=42, probe_range=PROBE_RANGE, build_range=BUILD_RANGE,
                                   group_range=GROUP_RANGE)
    else:
        log = runlog.open_log(a.runlog) if a.runlog else None
        Xtr, Ttr, ptr = explore(a.train, seed=42, reps=a.reps, out=os.path.join(a.out, "partial_train.npz"), log=log)
        Xho, Tho, pho = explore(a.heldout, seed=4242, reps=a.reps, out=os.path.join(a.out, "partial_heldout.npz"),
                                log=log)
        np.savez(os.path.join(a.out, "timings.npz"), Xtr=Xtr, Ttr=Ttr, Xho=Xho, Tho=Tho)
    t0 = time.time()
    net, mu, sd, losses = selector.train(Xtr, Ttr[:, :selector.N_VARIANTS], epochs=a.epochs)
    train_s = time.time() - t0
    print(f"trained {a.epochs} epochs in {train_s:.1f}s: loss {losses[0]:.3f} -> {losses[-1]:.3f}", flush=True)
    np.savez(os.path.join(a.out, "model.npz"), params=net.params(), mu=mu, sd=sd, losses=np.array(losses))
    rep = evaluate(net, mu, sd, Xho, Tho, pho, fixed=fixed_blend(Ttr, Tho))
    rep["train"] = {"pipelines": int(len(Xtr)), "epochs": a.epochs, "seconds": train_s,
                    "loss_first": losses[0], "loss_last": losses[-1]}
    rep["train_fit"] = evaluate(net, mu, sd, Xtr, Ttr, ptr)
    with open(os.path.join(a.out, "report.json"), "w") as f:
        json.dump(rep, f, indent=1)
    print(json.dumps(rep), flush=True)


if __name__ == "__main__":
    sys.exit(main())
