"""The benchmark configurations at their full sizes on the GPU, against the
oracle's vector(1024) restatement on the host: the star join with a
60,000,000-row fact table and full dimensions (uniform and Zipf 1.0 part
keys), and a batch of the batched workload's largest pipelines (10,000,000
probe rows, 2,000,000 build rows, up to 1,000,000 groups, key skew 1.5) in
one persistent launch, replayed from its graph; and a VOILA program run on
device columns."""
import os

import numpy as np
import pytest

import oracle
from vldb_voila_amd import datagen, relalg, runtime, workloads

import voila_programs as V

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else 8))


@pytest.fixture(scope="module", autouse=True)
def device():
    print("gpu full scale: init device 0", flush=True)
    runtime.init(0)


@pytest.mark.parametrize("theta", [0.0, 1.0])
def test_star_schema_full_size(theta):
    print(f"star 60M fact rows theta={theta}: generate", flush=True)
    fact, date, cust, supp, part = datagen.gen_star(60_000_000, zipf_part=theta, seed=42)
    print("star: oracle", flush=True)
    want = oracle.sorted_rows(oracle.run_pipeline(
        workloads.star_pipeline(fact, date, cust, supp, part).desc(), threads=THREADS)[0])
    assert len(want) == 175
    print("star: GPU", flush=True)
    dev = {k: runtime.upload(v) for k, v in
           (("fact", fact), ("date", date), ("customer", cust), ("supplier", supp), ("part", part))}
    addrs = {k: runtime.addrs(v) for k, v in dev.items()}
    q = runtime.Query(workloads.star_pipeline(fact, date, cust, supp, part, addrs).desc())
    for rep in range(2):
        q.run()
        assert oracle.sorted_rows(q.result()) == want
    # size-independent: the counts add up to the fact rows every join keeps
    total = sum(int(r.split("|")[4]) for r in want)
    c_of = np.empty(len(cust["c_key"]) + 1, np.int32)
    c_of[cust["c_key"]] = cust["c_region"]
    s_of = np.empty(len(supp["s_key"]) + 1, np.int32)
    s_of[supp["s_key"]] = supp["s_region"]
    keep = (c_of[fact["f_cust"]] == workloads.STAR_REGION) & (s_of[fact["f_supp"]] == workloads.STAR_REGION)
    assert total == int(keep.sum())


def test_batch_of_the_largest_pipelines():
    params = [
        {"probe_rows": 10_000_000, "build_rows": 2_000_000, "selectivity": 0.5, "groups": 1_000_000,
         "theta": 1.5},
        {"probe_rows": 10_000_000, "build_rows": 1_000_000, "selectivity": 0.99, "groups": 1_000_000,
         "theta": 0.0},
        {"probe_rows": 5_000_000, "build_rows": 2_000_000, "selectivity": 0.2, "groups": 1000, "theta": 1.0},
        {"probe_rows": 10_000_000, "build_rows": 1000, "selectivity": 0.7, "groups": 1, "theta": 1.5},
    ]
    descs, devs, wants = [], [], []
    for i, prm in enumerate(params):
        print(f"batch: pipeline {i} {prm}", flush=True)
        b, p = datagen.gen_batch_pipeline(prm, seed=42, index=500 + i)
        wants.append(oracle.sorted_rows(oracle.run_pipeline(workloads.batch_pipeline(prm, b, p).desc(),
                                                            threads=THREADS)[0]))
        db, dp = runtime.upload(b), runtime.upload(p)
        devs.append((db, dp, b, p))
        descs.append(workloads.batch_pipeline(prm, b, p, runtime.addrs(db), runtime.addrs(dp)).desc())
    assert len(wants[0]) > 100_000
    print("batch: GPU", flush=True)
    batch = runtime.Batch(descs)
    s = runtime.Stream()
    batch.run(s)
    for i in range(len(descs)):
        assert oracle.sorted_rows(batch.result(i, s)) == wants[i], i
    for rep in range(2):
        batch.run_graph(s)
        s.synchronize()
        for i in range(len(descs)):
            assert oracle.sorted_rows(batch.result(i, s)) == wants[i], (rep, i)
    batch.free()


def test_voila_program_on_device_columns():
    """The headline join written as the translator's VOILA program, lowered
    and run on the GPU."""
    print("VOILA program: join + aggregate 1.5M x 6M", flush=True)
    b, p = datagen.gen_join_aggregate(1_500_000, 6_000_000, seed=42)
    db_, dp_ = runtime.upload(b), runtime.upload(p)
    db = relalg.Database().add_table("probe", p, runtime.addrs(dp_)).add_table("build", b, runtime.addrs(db_))
    plan = ("hashaggr", "hash",
            ("hashjoin", "joinn",
             ("select", ("le", "date", workloads.JOIN_DATE_CUTOFF), ("scan", "probe", ["fk", "date", "price"])),
             ["fk"], ("scan", "build", ["key", "segment"]), ["key"], ["segment"]),
            ["segment"], [("sum", "price")])
    q = db.query_program(V.program(plan))
    q.run()
    want = oracle.sorted_rows(oracle.bruteforce(workloads.join_aggregate_pipeline(b, p).desc()))
    assert oracle.sorted_rows(q.result()) == want
