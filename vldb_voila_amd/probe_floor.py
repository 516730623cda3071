"""What bounds the headline join's probe kernel: the same 6,000,000 probe rows
(fk, date, price) run through three pipelines whose generated probe kernels
differ in one thing each, timed by their own start and end stamps
(hipExtModuleLaunchKernel events) over rotating input copies:

  stream    select date <= cutoff; sum(price), sum(fk) -- the three columns
            streamed, no table lookup
  l1_table  the join against a 64-key build table (fk folded onto 1..64):
            every lookup hits the 64-byte direct table in L1
  join      the headline join (1.5 M keys, a 1.5 MB direct table: every
            lookup an L1 miss served by L2)

    python3 -m vldb_voila_amd.probe_floor [--steps 40]

Prints one JSON line: each pipeline's mean, min and median probe-kernel time.
The difference join - stream is the cost of the random lookups; l1_table -
stream the cost of the join's instructions with every lookup an L1 hit.
"""
from __future__ import annotations

import argparse
import json

import numpy as np

from . import abi, datagen, runtime, workloads
from .pipeline import Pipeline, col, table_from_arrays


def _stream(probe, addrs=None):
    s = table_from_arrays("probe", probe, addrs)
    p = Pipeline(s)
    p.select("date", "<=", workloads.JOIN_DATE_CUTOFF)
    p.groupby([], [("sum", col("price")), ("sum", col("fk"))])
    return p


def _l1_table(build64, probe64, baddrs=None, paddrs=None):
    return workloads.join_aggregate_pipeline(build64, probe64, baddrs, paddrs)


def time_probe(make, tables, steps, copies, **cfg):
    """Mean, min, median of the probe kernel's stamped duration (ms)."""
    stream = runtime.Stream()
    queries, bufs = [], []
    for _ in range(copies):
        dev = [runtime.upload(t) for t in tables]
        bufs.append(dev)
        queries.append(runtime.Query(make(*[runtime.addrs(d) for d in dev]).desc(),
                                     runtime.make_config(**cfg) if cfg else None))
    for q in queries:
        q.run(stream)
    ev = [(runtime.Event(), runtime.Event()) for _ in range(steps)]
    for i in range(steps):
        q = queries[i % copies]
        q.run_phase(0, stream)
        q.run_phase(1, stream)
        q.probe_events(*ev[i])
        q.run_phase(2, stream)
        q.run_phase(3, stream)
    stream.synchronize()
    for q in queries:
        q.probe_events(None, None)
    ms = sorted(a.elapsed_ms(b) for a, b in ev)
    rows = queries[0].result(stream)
    for q in queries:
        q.free()
    return {"mean_ms": float(np.mean(ms)), "min_ms": ms[0], "median_ms": ms[len(ms) // 2]}, rows


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--copies", type=int, default=8)
    args = ap.parse_args()
    runtime.init(0)
    build, probe = datagen.gen_join_aggregate(1_500_000, 6_000_000, seed=42)
    build64 = {k: v[:64].copy() for k, v in build.items()}
    build64["key"] = np.arange(1, 65, dtype=np.int32)
    probe64 = dict(probe)
    probe64["fk"] = ((probe["fk"].astype(np.int64) - 1) % 64 + 1).astype(np.int32)
    out = {}
    out["stream"], _ = time_probe(lambda a: _stream(probe, a), [probe], args.steps, args.copies)
    out["stream_reg"], _ = time_probe(lambda a: _stream(probe, a), [probe], args.steps, args.copies,
                                      group_mode=abi.VH_GROUP_REG)
    out["l1_table"], _ = time_probe(lambda a, b: _l1_table(build64, probe64, a, b), [build64, probe64],
                                    args.steps, args.copies)
    out["join"], _ = time_probe(lambda a, b: workloads.join_aggregate_pipeline(build, probe, a, b),
                                [build, probe], args.steps, args.copies)
    out["probe_rows"] = 6_000_000
    out["column_bytes"] = 6_000_000 * 16
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
