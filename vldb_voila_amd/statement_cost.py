"""What the statement-at-a-time path costs against the fused pipeline, on the
same device-resident inputs:

  groupby  the HashAggr statement loop (statements.groupby_statements: hash,
           bucket_lookup, check, aggr, bucket_next, selunion, bucket_insert,
           scatter) over 6,000,000 rows of the filter workload's table,
           group by (flag, status): sum(qty), sum(price), count; against the
           same group-by as one fused pipeline query
  join     the HashJoin statement loop (statements.join_statements: build
           with write_pos / write / bucket_build, probe with bucket_lookup,
           check, gather, bucket_next) of the headline join's 1,500,000 build
           and 6,000,000 probe rows, payload segment; against the headline
           pipeline (the same join, its date filter and sum(price) by
           segment) as one fused query

    python3 -m vldb_voila_amd.statement_cost [--reps 3]

Prints one JSON line: per loop its wall time (ms, mean and min over the
repetitions, the loop synchronises with the host after every selection to
size the next statement), the number of those host round trips, and the
fused query's time per run (HIP events around back-to-back runs).  The
statement group-by's rows are checked against the fused query's, the
statement join's matches against the build table (one match per probe row,
its key's segment).
"""
from __future__ import annotations

import argparse
import json
import time

import numpy as np

from . import abi, datagen, runtime, statements, workloads
from .pipeline import Pipeline, col, table_from_arrays


def _timed_runs(q, stream, reps):
    q.run(stream)
    stream.synchronize()
    e0, e1 = runtime.Event(), runtime.Event()
    e0.record(stream)
    for _ in range(reps):
        q.run(stream)
    e1.record(stream)
    stream.synchronize()
    return e0.elapsed_ms(e1) / reps


def _wall(fn, reps, trips):
    fn()   # warm-up
    times, n = [], 0
    for _ in range(reps):
        trips[0] = 0
        t0 = time.perf_counter()
        out = fn()
        times.append((time.perf_counter() - t0) * 1e3)
        n = trips[0]
    return out, {"mean_ms": float(np.mean(times)), "min_ms": float(min(times)), "host_round_trips": n}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--reps", type=int, default=3)
    args = ap.parse_args()
    runtime.init(0)
    stream = runtime.Stream()
    trips = [0]
    count = statements._count

    def counted(d):
        trips[0] += 1
        return count(d)
    statements._count = counted
    out = {}

    # group-by
    rows = 6_000_000
    t = datagen.gen_filter_groupby(rows, seed=42)
    dt = runtime.upload(t)
    keys = [(abi.VH_I32, dt["flag"].ptr), (abi.VH_I32, dt["status"].ptr)]
    vals = [(abi.VH_I64, dt["qty"].ptr), (abi.VH_I64, dt["price"].ptr), (abi.VH_I64, None)]
    (tbl, n), out["groupby_statements"] = _wall(lambda: statements.groupby_statements(keys, vals, rows), args.reps,
                                                 trips)
    cols = [tbl.column(c, n) for c in range(5)]
    got = sorted("|".join(str(cols[c][i]) for c in range(5)) for i in range(n))
    p = Pipeline(table_from_arrays("t", t, runtime.addrs(dt)))
    p.groupby(["flag", "status"], [("sum", col("qty")), ("sum", col("price")), ("count",)])
    q = runtime.Query(p.desc())
    out["groupby_fused_ms"] = _timed_runs(q, stream, 20)
    want = sorted(line for line in q.result(stream).splitlines() if line)
    q.free()
    if got != want:
        raise SystemExit("statement group-by differs from the fused query")

    # join
    build, probe = datagen.gen_join_aggregate(1_500_000, 6_000_000, seed=42)
    db, dp = runtime.upload(build), runtime.upload(probe)
    (pos, pays), out["join_statements"] = _wall(
        lambda: statements.join_statements((abi.VH_I32, db["key"].ptr), [(abi.VH_I32, db["segment"].ptr)],
                                           1_500_000, (abi.VH_I32, dp["fk"].ptr), 6_000_000, abi.VH_JOIN01),
        args.reps, trips)
    q = runtime.Query(workloads.join_aggregate_pipeline(build, probe, runtime.addrs(db), runtime.addrs(dp)).desc())
    out["join_fused_ms"] = _timed_runs(q, stream, 20)
    q.free()
    # every probe key has one build row: the statement join's matches are
    # the probe rows, each with its key's segment
    seg = np.empty(1_500_001, dtype=np.int64)
    seg[build["key"]] = build["segment"]
    order = np.argsort(pos, kind="stable")
    if len(pos) != 6_000_000 or not np.array_equal(np.asarray(pays[0])[order], seg[probe["fk"][np.asarray(pos)[order]]]):
        raise SystemExit("statement join differs from the build table")
    out["groupby_rows"] = rows
    out["join_rows"] = [1_500_000, 6_000_000]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
