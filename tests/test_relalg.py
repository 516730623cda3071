"""Relational front end (CPU): relalg plans in S-expression form lower to the
same pipelines the workload builders describe; the oracle gives identical
answers for both; unsupported shapes are rejected with an error."""
import numpy as np
import pytest

import oracle
from vldb_voila_amd import abi, datagen, relalg, runtime, workloads


def test_filter_groupby_plan_matches_builder():
    t = datagen.gen_filter_groupby(30_000, seed=3)
    db = relalg.Database().add_table("lineitem", t)
    d = db.lower(relalg.filter_groupby_plan(workloads.FILTER_GROUPBY_CUTOFF))
    want = oracle.sorted_rows(oracle.bruteforce(workloads.filter_groupby_pipeline(t).desc()))
    assert oracle.sorted_rows(oracle.bruteforce(d)) == want
    assert oracle.sorted_rows(oracle.run_pipeline(d, threads=2)[0]) == want
    text = db.explain(relalg.filter_groupby_plan(workloads.FILTER_GROUPBY_CUTOFF))
    assert "hashaggr" in text and "vec_seltrue" in text


@pytest.mark.parametrize("variant", ["join01", "joinn"])
def test_join_plan_matches_builder(variant):
    b, p = datagen.gen_join_aggregate(5000, 20000, seed=4)
    db = relalg.Database().add_table("probe", p).add_table("build", b)
    d = db.lower(relalg.join_aggregate_plan(workloads.JOIN_DATE_CUTOFF, variant))
    assert d.joins[0].variant == (abi.VH_JOIN01 if variant == "join01" else abi.VH_JOINN)
    want = oracle.sorted_rows(oracle.bruteforce(workloads.join_aggregate_pipeline(b, p).desc()))
    assert oracle.sorted_rows(oracle.bruteforce(d)) == want


def test_star_plan_matches_builder():
    f, dd, c, s, p = datagen.gen_star(200_000, zipf_part=1.0, dims_scale=0.05, seed=5)
    db = relalg.Database()
    for name, t in (("fact", f), ("date", dd), ("customer", c), ("supplier", s), ("part", p)):
        db.add_table(name, t)
    d = db.lower(relalg.star_plan(workloads.STAR_REGION))
    want = oracle.sorted_rows(oracle.bruteforce(workloads.star_pipeline(f, dd, c, s, p).desc()))
    got = oracle.sorted_rows(oracle.bruteforce(d))
    assert got == want and len(want) <= 175


def test_project_assignments_and_global_aggregate():
    t = datagen.gen_filter_groupby(10_000, seed=6)
    db = relalg.Database().add_table("l", t)
    plan = """(hashaggr global (keys) (aggrs (sum disc_price) (count))
      (project (assign disc_price (* price (- 100 disc)))
        (select (and (>= flag 1) (< status 1)) (scan l flag status price disc))))"""
    d = db.lower(plan)
    m = (t["flag"] >= 1) & (t["status"] < 1)
    want = [f"{int((t['price'][m].astype(object) * (100 - t['disc'][m])).sum())}|{int(m.sum())}"]
    assert oracle.sorted_rows(oracle.bruteforce(d)) == want


@pytest.mark.parametrize("plan", [
    "(hashaggr hash (keys x) (aggrs (count)) (scan nosuch x))",
    "(hashaggr hash (keys flag) (aggrs (max price)) (scan l flag price))",
    "(scan l flag)",
    "(hashaggr hash (keys flag) (aggrs (count)) (scan l flag)",
])
def test_rejected_plans(plan):
    t = datagen.gen_filter_groupby(100, seed=1)
    db = relalg.Database().add_table("l", t)
    with pytest.raises(runtime.VoilaHipError):
        db.lower(plan)


def test_multi_column_key_join_plan():
    """A join on two key columns with two payloads (the translator's key and
    payload lists, relalg_translator.cpp:812-1115) lowers to the same
    pipeline as the builder's multi-column join."""
    from vldb_voila_amd.pipeline import Pipeline, col, table_from_arrays
    rng = np.random.default_rng(12)
    ps = {"ps_partkey": rng.integers(1, 200, 3000).astype(np.int32),
          "ps_suppkey": rng.integers(1, 40, 3000).astype(np.int32),
          "ps_cost": rng.integers(100, 1000, 3000).astype(np.int64),
          "ps_qty": rng.integers(1, 10, 3000).astype(np.int32)}
    li = {"l_partkey": rng.integers(1, 210, 20000).astype(np.int32),
          "l_suppkey": rng.integers(1, 42, 20000).astype(np.int32),
          "l_qty": rng.integers(1, 50, 20000).astype(np.int64)}
    db = relalg.Database().add_table("lineitem", li).add_table("partsupp", ps)
    plan = """(hashaggr hash (keys ps_qty) (aggrs (sum (* l_qty ps_cost)) (count))
      (hashjoin joinn (scan lineitem l_partkey l_suppkey l_qty) (keys l_partkey l_suppkey)
        (payl) (scan partsupp ps_partkey ps_suppkey ps_cost ps_qty) (keys ps_partkey ps_suppkey)
        (payl ps_cost ps_qty)))"""
    d = db.lower(plan)
    assert d.joins[0].nkeys == 2 and d.joins[0].npayload == 2
    p = Pipeline(table_from_arrays("lineitem", li))
    p.join(table_from_arrays("partsupp", ps), ["ps_partkey", "ps_suppkey"], ["l_partkey", "l_suppkey"],
           payload=["ps_cost", "ps_qty"], names=["ps_cost", "ps_qty"])
    p.groupby(["ps_qty"], [("sum", col("l_qty") * col("ps_cost")), ("count",)])
    want = oracle.sorted_rows(oracle.bruteforce(p.desc()))
    assert len(want) == 9
    assert oracle.sorted_rows(oracle.bruteforce(d)) == want
    assert oracle.sorted_rows(oracle.run_pipeline(d, threads=2)[0]) == want
