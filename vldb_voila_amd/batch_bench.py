"""Batched uneven pipelines: one persistent hipGraph-replayed launch against
the same pipelines run one after the other with plain launches.

    python -m vldb_voila_amd.batch_bench [--count 256] [--scale 1.0] [--reps 5]

256 select -> join -> group-by pipelines (seed 42): probe rows log-uniform in
[1e4, 1e7] x scale, build rows in [1e3, 2e6] x scale, selectivity U[0.01,
0.99], groups log-uniform in [1, 1e6], key skew theta in {0, .5, 1, 1.5},
int64 sums.  Every pipeline's batched result must equal its sequential
result; the smallest ones are also checked against the CPU oracle.  Prints
one JSON line.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import datagen, runtime, workloads


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--count", type=int, default=256)
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--oracle-checks", type=int, default=8)
    a = ap.parse_args(argv)
    runtime.init(0)
    t0 = time.time()
    params = datagen.batch_params(a.count, seed=42, probe_range=(1e4 * a.scale, 1e7 * a.scale),
                                  build_range=(max(1e3 * a.scale, 10), 2e6 * a.scale))
    descs, devs, data = [], [], []
    rows = 0
    for i, prm in enumerate(params):
        b, p = datagen.gen_batch_pipeline(prm, seed=42, index=i)
        db, dp = runtime.upload(b), runtime.upload(p)
        devs.append((db, dp))
        descs.append(workloads.batch_pipeline(prm, b, p, runtime.addrs(db), runtime.addrs(dp)).desc())
        rows += prm["probe_rows"] + prm["build_rows"]
        if a.oracle_checks:
            data.append((prm["probe_rows"], i, prm, b, p))
        else:
            del b, p
        if (i + 1) % 32 == 0:
            print(f"generated {i + 1}/{a.count} pipelines, {time.time() - t0:.0f}s", flush=True)
    stream = runtime.Stream()

    # sequential: one query per pipeline, plain launches (clear, build, probe)
    queries = [runtime.Query(d) for d in descs]
    for q in queries:
        q.run(stream)
    stream.synchronize()
    seq_rows = [q.result(stream) for q in queries]
    e0, e1 = runtime.Event(), runtime.Event()
    seq_ms = []
    for _ in range(a.reps):
        e0.record(stream)
        for q in queries:
            q.run(stream)
        e1.record(stream)
        stream.synchronize()
        seq_ms.append(e0.elapsed_ms(e1))
    print(f"sequential: {min(seq_ms):.3f} ms", flush=True)

    # batched: one persistent launch, captured into a hipGraph and replayed
    batch = runtime.Batch(descs)
    batch.run_graph(stream)
    stream.synchronize()
    bad = [i for i in range(a.count) if batch.result(i, stream) != seq_rows[i]
           and sorted(batch.result(i, stream).splitlines()) != sorted(seq_rows[i].splitlines())]
    bat_ms = []
    for _ in range(a.reps):
        e0.record(stream)
        batch.run_graph(stream)
        e1.record(stream)
        stream.synchronize()
        bat_ms.append(e0.elapsed_ms(e1))
    print(f"batched: {min(bat_ms):.3f} ms", flush=True)
    bad += [i for i in range(a.count) if sorted(batch.result(i, stream).splitlines())
            != sorted(seq_rows[i].splitlines())]

    oracle_ok = None
    if a.oracle_checks:
        import oracle
        data.sort(key=lambda x: x[0])
        oracle_ok = True
        for _, i, prm, b, p in data[:a.oracle_checks]:
            want = oracle.sorted_rows(oracle.bruteforce(workloads.batch_pipeline(prm, b, p).desc()))
            oracle_ok &= oracle.sorted_rows(batch.result(i, stream)) == want
    out = {
        "workload": f"{a.count} uneven select-join-group-by pipelines (seed 42, scale {a.scale})",
        "input_tuples": rows,
        "batched_ms": min(bat_ms), "sequential_ms": min(seq_ms),
        "speedup": min(seq_ms) / min(bat_ms),
        "batched_tuples_per_s": rows / (min(bat_ms) / 1e3),
        "sequential_tuples_per_s": rows / (min(seq_ms) / 1e3),
        "batch": batch.info(),
        "mismatches": sorted(set(bad)),
        "oracle_checked": a.oracle_checks, "oracle_ok": oracle_ok,
    }
    print(json.dumps(out), flush=True)
    return 0 if not bad and oracle_ok is not False else 1


if __name__ == "__main__":
    sys.exit(main())
