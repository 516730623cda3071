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
