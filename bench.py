"""Benchmark of the VOILA operator path on one MI355X.

    python bench.py [--workload join|filter|star|batch|contains] [--steps K] [--warmup W]

join (default, the headline configuration): build 1,500,000 rows (unique
  int32 keys 1..1.5M permuted, int32 nation U[0,25), int32 segment U[0,5));
  probe 6,000,000 rows (int32 fk uniform over the build keys, int32 date over
  2,526 days with a predicate keeping 50%, int64 price U[90000, 10500000));
  sum(price) grouped by the build side's segment.
filter: 6,000,000 rows, date predicate keeping ~96%, group by (flag,
  status): sum(qty), sum(price), sum(price * (100 - disc)) as int128, count.
star: fact 60,000,000 rows x date / customer (region 2) / supplier
  (region 2) / part, FKs uniform or (--zipf 1.0) Zipf on part; group by
  (c_nation, s_nation, year): sum(revenue - cost), count.
batch: 256 uneven select -> join -> group-by pipelines in one persistent
  launch replayed from a hipGraph (the same work as plain launches one
  pipeline after another is timed beside it).

All inputs are synthetic (seed 42, counter-based splitmix64).  A step is one
full run on device-resident inputs: clear, build every join table, the fused
select + probe + aggregate kernel, the merge of per-workgroup partials.
Inputs rotate over enough copies that the working set exceeds the 256 MB
Infinity Cache, so the figures are HBM figures.

Prints one JSON line: `roofline` is the probe kernel (HIP events its launch
stamps with the kernel's own start and end, over K more steps after the timed
ones, since stamping slows a step by a few microseconds); `cpu_baseline` times the
oracle's vector(1024) restatement on the host cores over a bounded sample,
and its rows double as the check of the GPU result (the run fails on a
difference).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# HBM bytes per launch of each workload's probe kernel, from rocprofv3
# FETCH_SIZE + WRITE_SIZE runs scaled by a copy of known size; each file
# gives the commands and the counter values
TRAFFIC_FILES = {"join": "traffic_join_1.5Mx6M.md", "join_sparse": "traffic_join_sparse_1.5Mx6M.md",
                 "filter": "traffic_filter_groupby_6M.md", "star": "traffic_star_60M.md",
                 "contains": "traffic_contains_10M.md", "batch": "kernels_batch_256.md"}
METRIC = "input tuples/s + % HBM peak; learned-selector vs default/exhaustive blend"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", choices=["join", "join_sparse", "filter", "star", "batch", "contains"],
                    default="join")
    ap.add_argument("--join-layout", type=int, default=0, help="vh_config.join_layout (0 = auto)")
    ap.add_argument("--part-shift", type=int, default=0, help="vh_config.part_shift (0 = auto, -1 = none)")
    ap.add_argument("--cfg", action="append", default=[],
                    help="another vh_config field as name=value (grid, unroll, block, merge, table_shift, ...)")
    ap.add_argument("--zipf", type=float, default=0.0, help="star: Zipf theta of the part FKs")
    ap.add_argument("--copies", type=int, default=0, help="input copies rotated (0 = auto)")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="budget of the CPU baseline sample")
    ap.add_argument("--cpu-threads", type=int, default=0, help="0 = every usable host core, at most 16")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline (and so the result check)")
    ap.add_argument("--tuned-layout-only", action="store_true",
                    help="with --tuned: take only the group-by strategy and table layout of each pick")
    ap.add_argument("--tuned", action="store_true",
                    help="batch: per-pipeline variants from the learned selector (autotuner below 0.5 "
                         "confidence), timed beside the default configuration")
    return ap.parse_args()


# ---------------------------------------------------------------------------
# workloads: host arrays, a pipeline factory over device addresses, and the
# probe kernel's algorithmic bytes per launch
# ---------------------------------------------------------------------------

def workload_join(datagen, workloads):
    build, probe = datagen.gen_join_aggregate(1_500_000, 6_000_000, seed=42)
    sel = int((probe["date"] <= workloads.JOIN_DATE_CUTOFF).sum())

    def make(addrs=None):
        a = addrs or [None, None]
        return workloads.join_aggregate_pipeline(build, probe, a[0], a[1])

    def probe_bytes(q):
        # fk 4 + date 4 + price 8 B per probe row, one 8-byte hash-table read
        # (key i32 + segment i32) per row that passes the date predicate
        return 6_000_000 * 16 + sel * 8
    # the whole step: build input 12 B per row (key, nation, segment), a
    # hash-table write of 8 B per build row, then the probe's bytes
    step_bytes = 1_500_000 * 12 + 1_500_000 * 8 + 6_000_000 * 16 + sel * 8
    return {"name": "hash join 1.5M build x 6M probe (50% date filter) + sum(price) by segment (5 groups)",
            "tables": [build, probe], "make": make, "tuples": 7_500_000, "probe_bytes": probe_bytes,
            "step_bytes": step_bytes, "config": {"build_rows": 1_500_000, "probe_rows": 6_000_000}}


def workload_join_sparse(datagen, workloads):
    build, probe = datagen.gen_join_sparse(1_500_000, 6_000_000, seed=42)
    sel = int((probe["date"] <= workloads.JOIN_DATE_CUTOFF).sum())

    def make(addrs=None):
        a = addrs or [None, None]
        return workloads.join_aggregate_pipeline(build, probe, a[0], a[1])

    W = {"name": "hash join 1.5M build (unique keys over [0, 2^31)) x 6M probe (50% date filter) + "
                 "sum(price) by segment (5 groups)",
         "tables": [build, probe], "make": make, "tuples": 7_500_000,
         "probe_bytes": lambda q: 6_000_000 * 16 + sel * 8,
         "step_bytes": 1_500_000 * 12 + 1_500_000 * 8 + 6_000_000 * 16 + sel * 8,
         "config": {"build_rows": 1_500_000, "probe_rows": 6_000_000, "key_range": "[0, 2^31)"}}
    return W


def workload_filter(datagen, workloads):
    t = datagen.gen_filter_groupby(6_000_000, seed=42)

    def make(addrs=None):
        return workloads.filter_groupby_pipeline(t, addrs[0] if addrs else None)
    # flag 4 + status 4 + qty 8 + price 8 + disc 8 + date 4 B per row
    return {"name": "filter (96%) + group by (flag, status): 3 sums (one int128) + count, 6M rows",
            "tables": [t], "make": make, "tuples": 6_000_000,
            "probe_bytes": lambda q: 6_000_000 * 36, "step_bytes": 6_000_000 * 36, "config": {"rows": 6_000_000}}


def workload_star(datagen, workloads, theta):
    fact, date, cust, supp, part = datagen.gen_star(60_000_000, zipf_part=theta, seed=42)
    dims = sum(len(next(iter(t.values()))) for t in (date, cust, supp, part))

    def make(addrs=None):
        a = addrs or [None] * 5
        return workloads.star_pipeline(fact, date, cust, supp, part,
                                       {"fact": a[0], "date": a[1], "customer": a[2], "supplier": a[3],
                                        "part": a[4]})
    # the four int32 FKs and two int64 measures of every fact row; the step
    # also reads each dimension's key and used columns (12 B per row)
    return {"name": f"star join: fact 60M x date, customer, supplier (region filters), part; "
                    f"175 groups; part FKs Zipf theta={theta}",
            "tables": [fact, date, cust, supp, part], "make": make, "tuples": 60_000_000 + dims,
            "probe_bytes": lambda q: 60_000_000 * 32, "step_bytes": 60_000_000 * 32 + dims * 12,
            "config": {"fact_rows": 60_000_000, "dimension_rows": dims, "zipf_theta": theta}}


def measured_traffic(workload, zipf):
    """Counter-derived HBM bytes per probe launch recorded under profiles/,
    or None (no record for this workload)."""
    name = TRAFFIC_FILES.get(workload)
    if name is None or (workload == "star" and zipf):
        return None
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return int(float(line.split(":")[1]))
    return None


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(every GPU box has one GPU); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.workload == "batch":
        return bench_batch(args)
    if args.workload == "contains":
        return bench_contains(args)

    import numpy as np

    from vldb_voila_amd import datagen, runtime, workloads

    if args.workload == "join":
        W = workload_join(datagen, workloads)
    elif args.workload == "join_sparse":
        W = workload_join_sparse(datagen, workloads)
    elif args.workload == "filter":
        W = workload_filter(datagen, workloads)
    else:
        W = workload_star(datagen, workloads, args.zipf)
    bytes_per_copy = sum(a.nbytes for t in W["tables"] for a in t.values())
    copies = args.copies or max(1, -(-768 * 2**20 // bytes_per_copy))

    runtime.init(0)
    stream = runtime.Stream()
    queries, bufs = [], []
    for c in range(copies):
        dev = [runtime.upload(t) for t in W["tables"]]
        bufs.append(dev)
        extra = {k: int(v) for k, v in (kv.split("=", 1) for kv in args.cfg)}
        queries.append(runtime.Query(W["make"]([runtime.addrs(d) for d in dev]).desc(),
                                     runtime.make_config(join_layout=args.join_layout,
                                                         part_shift=args.part_shift, **extra)))
    q0 = queries[0]
    info = q0.info()
    kernel, _ = q0.kernel()
    probe_bytes = W["probe_bytes"](q0)
    q0.run(stream)
    gpu_rows = sorted(q0.result(stream).splitlines())

    for i in range(args.warmup):
        queries[i % copies].run(stream)
    stream.synchronize()

    # timed region: K whole steps, plain runs
    K = args.steps
    t_start, t_stop = runtime.Event(), runtime.Event()
    stream.synchronize()
    wall0 = time.perf_counter()
    t_start.record(stream)
    for i in range(K):
        queries[i % copies].run(stream)
    t_stop.record(stream)
    stream.synchronize()
    wall = time.perf_counter() - wall0
    ms_per_step = t_start.elapsed_ms(t_stop) / K

    # the probe kernel's own duration: K more steps whose probe launches
    # stamp events with the kernel's start and end (stamping adds a few
    # microseconds per step, so these steps are not the ones timed above)
    ev = [(runtime.Event(), runtime.Event()) for _ in range(K)]
    for i in range(K):
        q = queries[i % copies]
        q.run_phase(0, stream)
        q.run_phase(1, stream)
        q.probe_events(*ev[i])
        q.run_phase(2, stream)
        q.run_phase(3, stream)
    stream.synchronize()
    for q in queries:
        q.probe_events(None, None)
    probe_ms = sorted(a.elapsed_ms(b) for a, b in ev)
    probe_avg_ms = float(np.mean(probe_ms))
    achieved = probe_bytes / (probe_avg_ms / 1e3) / 1e9

    line = {
        "metric": METRIC,
        "value": W["tuples"] / (ms_per_step / 1e3),
        "unit": "input tuples/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int64",
        "data": "synthetic (seed 42, counter-based splitmix64)",
        "config": dict({"workload": W["name"], "input_copies": copies, "group_mode": info["group_mode"],
                        "grid": info["grid"],
                        "join_tables": [dict(zip(("entry_bytes", "entries"), q0.join(j)))
                                        for j in range(q0.desc.njoins)],
                        "probe_kernel": "generated (hipRTC)" if kernel == 2 else "generic",
                        "radix_probe": "vh_generated_join" in q0.kernel()[1]}, **W["config"]),
        "roofline": {
            "bound": "hbm",
            "kernel": ("vh_generated" if kernel == 2 else "vh::pipeline_kernel") +
                      " (select + join probes + group-by, one launch)",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            # the records are of the default configuration
            "traffic": None if (args.cfg or args.part_shift or args.join_layout) else
            measured_traffic(args.workload, args.zipf),
            "avg_kernel_ms": probe_avg_ms,
            "min_kernel_ms": probe_ms[0],
            "median_kernel_ms": probe_ms[len(probe_ms) // 2],
            "algorithmic_bytes": probe_bytes,
            "step_algorithmic_bytes": W["step_bytes"],
            "step_achieved": W["step_bytes"] / (ms_per_step / 1e3) / 1e9,
            "step_frac": W["step_bytes"] / (ms_per_step / 1e3) / 1e9 / HBM_PEAK_GBS,
        },
        "wall_s": wall,
    }
    if not args.no_cpu:
        base, cpu_rows = cpu_baseline(W["make"]().desc(), W["tuples"], args)
        line["cpu_baseline"] = base
        if cpu_rows != gpu_rows:
            print(json.dumps({"error": "GPU result differs from the CPU oracle", "gpu": gpu_rows[:8],
                              "cpu": cpu_rows[:8]}), flush=True)
            sys.exit(1)
    print(json.dumps(line), flush=True)


def host_cpu():
    """(CPU model, nproc, cores this process may use)."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    allowed = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    return model, os.cpu_count() or 1, allowed


def cpu_threads(args):
    """Every core this process may use, at most 16 (a GPU box's CPU share
    for one GPU, whatever nproc reports there)."""
    _, _, allowed = host_cpu()
    return args.cpu_threads or max(1, min(allowed, 16))


def cpu_baseline(desc, tuples, args):
    """The oracle's vector(1024) restatement on the host cores: the full
    workload, 5 hot runs (min and avg, as compiler.cpp:576-588 reports them)
    on every usable core, and the same at one thread while the budget
    lasts.  Returns (baseline record, its sorted result rows)."""
    import oracle

    model, nproc, allowed = host_cpu()
    threads = cpu_threads(args)

    def hot_runs(t, budget):
        times, text = [], ""
        t0 = time.perf_counter()
        oracle.run_pipeline(desc, threads=t)          # cold run, not counted
        while len(times) < 5 and (len(times) < 1 or time.perf_counter() - t0 < budget):
            text, ns = oracle.run_pipeline(desc, threads=t)
            times.append(ns / 1e9)
        return times, text

    times, text = hot_runs(threads, args.cpu_seconds)
    t1, _ = hot_runs(1, args.cpu_seconds)
    best, best1 = min(times), min(t1)
    rec = {"value": tuples / best, "unit": "input tuples/s", "cores": threads, "kind": "port",
           "cpu_model": model, "nproc": nproc, "usable_cores": allowed,
           "avg_value": tuples / (sum(times) / len(times)),
           "t1_value": tuples / best1, "t1_avg_value": tuples / (sum(t1) / len(t1)),
           "sample": f"full workload; vector(1024) restatement, {len(times)} hot runs at {threads} threads "
                     f"(min {best * 1e3:.1f} ms, avg {sum(times) / len(times) * 1e3:.1f} ms) and {len(t1)} at "
                     f"1 thread (min {best1 * 1e3:.1f} ms, avg {sum(t1) / len(t1) * 1e3:.1f} ms); "
                     f"{model}, nproc {nproc}, {allowed} usable, pool capped at 16"}
    return rec, sorted(text.splitlines())


def bench_contains(args):
    """vec_contains over a p_name-like varchar(55) column: 10M rows of
    56-byte slots, needle 'green' (TPC-H Q9's predicate); two copies rotate
    so the 560 MB column is read from HBM.  One kernel per step."""
    import numpy as np

    from vldb_voila_amd import datagen, runtime, varchar

    rows, max_len, needle = 10_000_000, 55, b"green"
    col = datagen.gen_part_names(rows, seed=42)
    runtime.init(0)
    stream = runtime.Stream()
    copies = [runtime.DeviceBuffer.from_numpy(col) for _ in range(2)]
    out = runtime.DeviceBuffer(rows)
    for i in range(max(args.warmup, 1)):
        varchar.contains(copies[i % 2].ptr, rows, max_len, needle, out=out, stream=stream)
    stream.synchronize()
    gpu = out.to_numpy(np.uint8, rows)
    K = args.steps
    t0, t1 = runtime.Event(), runtime.Event()
    wall0 = time.perf_counter()
    t0.record(stream)
    for i in range(K):
        varchar.contains(copies[i % 2].ptr, rows, max_len, needle, out=out, stream=stream)
    t1.record(stream)
    stream.synchronize()
    wall = time.perf_counter() - wall0
    ms = t0.elapsed_ms(t1) / K
    alg = rows * (max_len + 1) + rows          # every slot read once, one result byte written
    achieved = alg / (ms / 1e3) / 1e9
    line = {
        "metric": METRIC, "value": rows / (ms / 1e3), "unit": "input tuples/s", "n_gpus": 1, "steps": K,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u8", "data": "synthetic p_name-like strings (seed 42)",
        "config": {"workload": "vec_contains(p_name, 'green'): 10M rows of varchar(55) slots", "rows": rows},
        "roofline": {"bound": "hbm", "kernel": "vh::varchar_kernel (contains)", "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic("contains", 0.0), "avg_kernel_ms": ms, "algorithmic_bytes": alg},
        "wall_s": wall,
    }
    if not args.no_cpu:
        import oracle
        sample = 2_000_000
        want = np.zeros(sample, dtype=np.uint8)
        times = []
        t_start = time.perf_counter()
        while len(times) < 5 and (len(times) < 2 or time.perf_counter() - t_start < args.cpu_seconds):
            c0 = time.perf_counter()
            oracle.lib().vo_vec_contains(None, sample, want.ctypes.data, col.ctypes.data, max_len, needle,
                                         len(needle))
            times.append(time.perf_counter() - c0)
        best = min(times)
        line["cpu_baseline"] = {"value": sample / best, "unit": "input tuples/s", "cores": 1, "kind": "port",
                                "sample": f"first 2M rows, memmem per row (varchar::contains), {len(times)} "
                                          f"runs, min {best * 1e3:.1f} ms"}
        if not np.array_equal(gpu[:sample], want):
            print(json.dumps({"error": "GPU contains differs from the CPU oracle"}), flush=True)
            sys.exit(1)
    print(json.dumps(line), flush=True)


def bench_batch(args):
    """256 uneven pipelines: one hipGraph replay of the persistent launches
    against plain launches one pipeline after another."""
    from vldb_voila_amd import datagen, runtime, workloads

    runtime.init(0)
    stream = runtime.Stream()
    params = datagen.batch_params(256, seed=42)
    descs, devs, hosts = [], [], []
    tuples = 0
    for i, prm in enumerate(params):
        b, p = datagen.gen_batch_pipeline(prm, seed=42, index=i)
        db, dp = runtime.upload(b), runtime.upload(p)
        devs.append((db, dp))
        descs.append(workloads.batch_pipeline(prm, b, p, runtime.addrs(db), runtime.addrs(dp)).desc())
        hosts.append((prm, b, p))
        tuples += prm["probe_rows"] + prm["build_rows"]
    # the same pipelines as plain launches one after another: the generic
    # kernel the batch runs, and the default (generated) kernels
    queries = [runtime.Query(d, runtime.make_config(kernel=1)) for d in descs]
    queries_gen = [runtime.Query(d) for d in descs]
    batch = runtime.Batch(descs)
    tuned = None
    if args.tuned:
        # per-pipeline variants: the selector's pick, or the autotuner's
        # exhaustive timing when the network is not confident
        from vldb_voila_amd import selector
        net, mu, sd = selector.load_model()
        cfgs, fallbacks = [], 0
        t_tune = time.perf_counter()
        for i, d in enumerate(descs):
            cfg, v, fb = selector.pick_config(net, mu, sd, d, hosts[i][2], "fk", stream=stream, batch=True)
            if args.tuned_layout_only:
                # the variant's group-by strategy and table layout; the batch's
                # own launch shape (work items, rows per lane)
                cfg.tile_shift, cfg.unroll, cfg.block, cfg.grid = 0, 0, 0, 0
            cfgs.append(cfg)
            fallbacks += fb
        t_tune = time.perf_counter() - t_tune
        tuned_batch = runtime.Batch(descs, cfgs)
        tuned_batch.run_graph(stream)
        stream.synchronize()
        tuned = (tuned_batch, fallbacks, t_tune)
    for _ in range(max(args.warmup, 1)):
        batch.run_graph(stream)
        for q in queries:
            q.run(stream)
    stream.synchronize()
    gpu_rows = [sorted(batch.result(i, stream).splitlines()) for i in range(len(descs))]
    seq_rows = [sorted(q.result(stream).splitlines()) for q in queries]
    if gpu_rows != seq_rows:
        print(json.dumps({"error": "batched and sequential results differ"}), flush=True)
        sys.exit(1)
    if tuned and [sorted(tuned[0].result(i, stream).splitlines()) for i in range(len(descs))] != gpu_rows:
        print(json.dumps({"error": "the tuned batch's results differ"}), flush=True)
        sys.exit(1)
    for q in queries_gen:
        q.run(stream)
    stream.synchronize()
    if [sorted(q.result(stream).splitlines()) for q in queries_gen] != seq_rows:
        print(json.dumps({"error": "generated and generic kernels differ"}), flush=True)
        sys.exit(1)
    e0, e1, e2, e3 = runtime.Event(), runtime.Event(), runtime.Event(), runtime.Event()
    K = args.steps
    wall0 = time.perf_counter()
    e0.record(stream)
    for _ in range(K):
        batch.run_graph(stream)
    e1.record(stream)
    wall = time.perf_counter() - wall0
    for _ in range(K):
        for q in queries:
            q.run(stream)
    e2.record(stream)
    for _ in range(K):
        for q in queries_gen:
            q.run(stream)
    e3.record(stream)
    stream.synchronize()
    batched_ms = e0.elapsed_ms(e1) / K
    seq_ms = e1.elapsed_ms(e2) / K
    seq_gen_ms = e2.elapsed_ms(e3) / K
    tuned_info = None
    if tuned:
        t0e, t1e = runtime.Event(), runtime.Event()
        t0e.record(stream)
        for _ in range(K):
            tuned[0].run_graph(stream)
        t1e.record(stream)
        stream.synchronize()
        tuned_ms = t0e.elapsed_ms(t1e) / K
        tuned_info = {"ms": tuned_ms, "speedup_vs_default": batched_ms / tuned_ms,
                      "autotuner_fallbacks": tuned[1], "tuning_seconds": tuned[2],
                      "selector": "vldb_voila_amd/data/selector_model.npz"}
    binfo = batch.info()
    # every probe row's columns (fk 4 + a 4 + g 4 + v 8 + w 8 B) and every
    # build row's key and payload (8 B)
    alg = sum(prm["probe_rows"] * 28 + prm["build_rows"] * 8 for prm, _, _ in hosts)
    line = {
        "metric": METRIC, "value": tuples / (batched_ms / 1e3), "unit": "input tuples/s", "n_gpus": 1,
        "steps": K, "warmup": args.warmup, "ms_per_step": batched_ms, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "int64",
        "data": "synthetic (seed 42, counter-based splitmix64)",
        "config": {"workload": "256 uneven select-join-group-by pipelines, one hipGraph of persistent launches",
                   "input_tuples": tuples, "batch": binfo, "sequential_ms": seq_ms,
                   "speedup_vs_sequential": seq_ms / batched_ms,
                   "sequential_generated_ms": seq_gen_ms,
                   "speedup_vs_sequential_generated": seq_gen_ms / batched_ms,
                   "tuned": tuned_info},
        "roofline": {"bound": "hbm", "kernel": "whole batch (clear + builds + probes)",
                     "achieved": alg / (batched_ms / 1e3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": alg / (batched_ms / 1e3) / 1e9 / HBM_PEAK_GBS,
                     "traffic": measured_traffic("batch", 0.0),
                     "avg_kernel_ms": batched_ms, "algorithmic_bytes": alg},
        "wall_s": wall,
    }
    if not args.no_cpu:
        # a bounded sample: the smallest pipelines until the budget is spent;
        # each one's rows are also the check of the batched GPU result
        import oracle
        threads = cpu_threads(args)
        order = sorted(range(len(hosts)), key=lambda i: hosts[i][0]["probe_rows"])
        done, ns_total, t_sample = 0, 0, 0
        t0 = time.perf_counter()
        for i in order:
            if done and time.perf_counter() - t0 > args.cpu_seconds:
                break
            prm, b, p = hosts[i]
            text, ns = oracle.run_pipeline(workloads.batch_pipeline(prm, b, p).desc(), threads=threads)
            if sorted(text.splitlines()) != gpu_rows[i]:
                print(json.dumps({"error": f"pipeline {i}: GPU result differs from the CPU oracle"}), flush=True)
                sys.exit(1)
            ns_total += ns
            t_sample += prm["probe_rows"] + prm["build_rows"]
            done += 1
        line["cpu_baseline"] = {"value": t_sample / (ns_total / 1e9), "unit": "input tuples/s",
                                "cores": threads, "kind": "port",
                                "sample": f"the {done} smallest of the 256 pipelines ({t_sample} tuples), "
                                          f"one run each of the vector(1024) restatement"}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
