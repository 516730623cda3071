"""Explorer, selector training and evaluation.

    python -m vldb_voila_amd.explore --train 4000 --heldout 500 --out DIR

1. Explorer: draws pipelines from the batched workload's distribution
   (probe / build rows and group counts log-uniform, selectivity uniform,
   key skew theta in {0, 0.5, 1, 1.5}; sizes shrunk to probe rows in
   [1e4, 5e5] and build rows in [1e3, 1e5] so 4,500 x 48 timings fit a few
   minutes), measures each one's features from its data (selector.features),
   and times all 48 kernel variants plus the built-in default of each on the
   GPU, and the wall time of that exhaustive search.  The reference's explorer does the same over CPU blends
   and stores the timings in SQLite (explorer.cpp:33-86, compiler.cpp:124-255).
2. Trains the selector MLP on the training pipelines with the HIP kernels.
3. Evaluates on the held-out pipelines (evaluate): the selector's choice
   (exhaustive fallback below 0.5 confidence, the search's measured time
   charged) against the exhaustive search on every pipeline, the built-in
   default heuristic, the untrained network, and the best fixed blend: the one variant (applicable
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
    S = np.zeros(count, dtype=np.float32)   # wall seconds of the exhaustive search
    stream = runtime.Stream()
    t0 = time.time()
    for i, prm in enumerate(params):
        prm = dict(prm)
        prm["groups"] = min(prm["groups"], prm["probe_rows"])
        b, p = datagen.gen_batch_pipeline(prm, seed=seed, index=i)
        db, dp = runtime.upload(b), runtime.upload(p)
        d = workloads.batch_pipeline(prm, b, p, runtime.addrs(db), runtime.addrs(dp)).desc()
        X[i] = selector.features(d, p, "fk")
        s0 = time.perf_counter()
        T[i] = selector.time_variants(d, reps=reps, stream=stream, configs=cfgs)
        S[i] = time.perf_counter() - s0
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
                np.savez(out, X=X[:i + 1], T=T[:i + 1], S=S[:i + 1])
    return X, T, S, params


def fixed_blend(T_train, T_eval):
    """Index of the variant that applies to every pipeline of both sets and
    has the lowest total time over the training pipelines."""
    vt, ve = T_train[:, :selector.N_VARIANTS], T_eval[:, :selector.N_VARIANTS]
    ok = (vt > 0).all(0) & (ve > 0).all(0)
    tot = np.where(vt > 0, vt, np.inf).sum(0)
    return int(np.argmin(np.where(ok, tot, np.inf)))


def policy(net, mu, sd, X, T):
    """A network's decision per pipeline: its most probable applicable
    variant, and whether that probability is below the confidence
    threshold (then the exhaustive search runs instead)."""
    var = T[:, :selector.N_VARIANTS]
    mask = (var > 0).astype(np.uint8)
    p = net.predict((X - mu) / sd, mask)
    pick = p.argmax(1)
    conf = p[np.arange(len(p)), pick]
    fallback = (conf < selector.CONFIDENCE) | (mask[np.arange(len(p)), pick] == 0)
    return pick, fallback


def evaluate(net, mu, sd, X, T, S, params, fixed=None, runs=100, baseline_net=None):
    """Price each policy on the held-out pipelines with the measured timings.
    Every pipeline runs `runs` times (a query executed repeatedly); a policy
    pays its variant's time per run plus, when it searches, the measured wall
    time of that pipeline's exhaustive 48-variant search (S, seconds):
      selector    the network's choice, or the search and its best when the
                  choice's probability is below 0.5;
      exhaustive  the search on every pipeline, then its best;
      fixed blend one variant for every pipeline (see fixed_blend);
      default     the built-in heuristic (the last timing column).
    baseline_net: another network (e.g. its random initialisation) priced
    the same way."""
    var = T[:, :selector.N_VARIANTS]
    default = T[:, selector.N_VARIANTS]
    best = np.where(var > 0, var, np.inf).min(1)
    search_ms = np.asarray(S, dtype=np.float64) * 1e3
    idx = np.arange(len(X))
    pick, fallback = policy(net, mu, sd, X, T)
    picked = np.where(var[idx, pick] > 0, var[idx, pick], best)
    chosen = np.where(fallback, best, picked)
    theta = np.array([q["theta"] for q in params])
    skew = theta >= 1.0
    R = float(runs)

    def total(per_run, searched):
        return float((R * per_run + np.where(searched, search_ms, 0.0)).sum())
    rep = {
        "heldout": int(len(X)),
        "runs_per_pipeline": runs,
        "fallback_fraction": float(fallback.mean()),
        "top1_accuracy": float(np.mean(pick == np.where(var > 0, var, np.inf).argmin(1))),
        "within_10pct_of_best": float(np.mean(chosen <= 1.1 * best)),
        "run_ms": {"selector": float(chosen.sum()), "selector_no_fallback": float(picked.sum()),
                   "exhaustive_best": float(best.sum()), "default": float(default.sum())},
        "search_ms_mean": float(search_ms.mean()),
        "policy_total_ms": {"selector": total(chosen, fallback), "exhaustive": total(best, np.ones(len(X), bool)),
                            "default": total(default, np.zeros(len(X), bool))},
        "selector_run_over_best": float(chosen.sum() / best.sum()),
        "default_run_over_selector": float(default.sum() / chosen.sum()),
        "default_run_over_selector_skewed": float(default[skew].sum() / chosen[skew].sum()) if skew.any() else None,
    }
    rep["selector_policy_over_exhaustive"] = rep["policy_total_ms"]["selector"] / rep["policy_total_ms"]["exhaustive"]
    if fixed is not None:
        f = var[:, fixed]
        rep["fixed_blend"] = {"variant": int(fixed), "run_ms": float(f.sum()),
                              "run_over_selector": float(f.sum() / chosen.sum()),
                              "run_over_selector_skewed": float(f[skew].sum() / chosen[skew].sum())
                              if skew.any() else None,
                              "run_over_best": float(f.sum() / best.sum()),
                              "policy_total_ms": total(f, np.zeros(len(X), bool))}
    if baseline_net is not None:
        bp, bf = policy(baseline_net, mu, sd, X, T)
        bpicked = np.where(var[idx, bp] > 0, var[idx, bp], best)
        bchosen = np.where(bf, best, bpicked)
        rep["baseline_net"] = {"run_ms_no_fallback": float(bpicked.sum()), "run_ms": float(bchosen.sum()),
                               "fallback_fraction": float(bf.mean()), "policy_total_ms": total(bchosen, bf)}
    return rep


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
        Xtr, Ttr, Str, Xho, Tho, Sho = z["Xtr"], z["Ttr"], z["Str"], z["Xho"], z["Tho"], z["Sho"]
        pho = datagen.batch_params(len(Xho), seed=4242, probe_range=PROBE_RANGE, build_range=BUILD_RANGE,
                                   group_range=GROUP_RANGE)
        ptr = datagen.batch_params(len(Xtr), seed=42, probe_range=PROBE_RANGE, build_range=BUILD_RANGE,
                                   group_range=GROUP_RANGE)
    else:
        log = runlog.open_log(a.runlog) if a.runlog else None
        Xtr, Ttr, Str, ptr = explore(a.train, seed=42, reps=a.reps, out=os.path.join(a.out, "partial_train.npz"),
                                     log=log)
        Xho, Tho, Sho, pho = explore(a.heldout, seed=4242, reps=a.reps,
                                     out=os.path.join(a.out, "partial_heldout.npz"), log=log)
        np.savez(os.path.join(a.out, "timings.npz"), Xtr=Xtr, Ttr=Ttr, Str=Str, Xho=Xho, Tho=Tho, Sho=Sho)
    t0 = time.time()
    net, mu, sd, losses = selector.train(Xtr, Ttr[:, :selector.N_VARIANTS], epochs=a.epochs)
    train_s = time.time() - t0
    print(f"trained {a.epochs} epochs in {train_s:.1f}s: loss {losses[0]:.3f} -> {losses[-1]:.3f}", flush=True)
    np.savez(os.path.join(a.out, "model.npz"), params=net.params(), mu=mu, sd=sd, losses=np.array(losses))
    untrained = selector.MLP(selector._he_init(np.random.default_rng(0)))
    rep = evaluate(net, mu, sd, Xho, Tho, Sho, pho, fixed=fixed_blend(Ttr, Tho), baseline_net=untrained)
    rep["train"] = {"pipelines": int(len(Xtr)), "epochs": a.epochs, "seconds": train_s,
                    "loss_first": losses[0], "loss_last": losses[-1]}
    rep["train_fit"] = evaluate(net, mu, sd, Xtr, Ttr, Str, ptr)
    with open(os.path.join(a.out, "report.json"), "w") as f:
        json.dump(rep, f, indent=1)
    print(json.dumps(rep), flush=True)


if __name__ == "__main__":
    sys.exit(main())
