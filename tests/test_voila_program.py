"""VOILA programs in (CPU): the statement sequences the reference's relalg
translator writes for scan, select, project, HashJoin (Join01 / JoinN, one
or two key columns, payload lists) and HashAggr (hash and global, with its
flush, read and re-aggregation pipelines), restated in tests/voila_programs,
lower through vh_program_lower to descriptors whose rows equal those of the
same plan lowered from its relalg text, and of the workload builders."""
import numpy as np
import pytest

import oracle
from vldb_voila_amd import abi, datagen, relalg, runtime, workloads

import voila_programs as V


def both(db, plan):
    dp = db.lower_program(V.program(plan))
    dr = db.lower(V.relalg_text(plan))
    return dp, dr


def rows(d):
    return oracle.sorted_rows(oracle.bruteforce(d))


def test_filter_groupby_program():
    t = datagen.gen_filter_groupby(30_000, seed=3)
    db = relalg.Database().add_table("lineitem", t)
    plan = ("hashaggr", "hash",
            ("select", ("le", "date", workloads.FILTER_GROUPBY_CUTOFF),
             ("scan", "lineitem", ["flag", "status", "qty", "price", "disc", "date"])),
            ["flag", "status"],
            [("sum", "qty"), ("sum", "price"), ("sum", ("*", "price", ("-", 100, "disc"))), ("count",)])
    dp, dr = both(db, plan)
    assert dp.nkeys == 2 and dp.naggrs == 4 and dp.probe.npreds == 1 and dp.njoins == 0
    want = oracle.sorted_rows(oracle.bruteforce(workloads.filter_groupby_pipeline(t).desc()))
    assert rows(dp) == rows(dr) == want
    assert oracle.sorted_rows(oracle.run_pipeline(dp, threads=2)[0]) == want


@pytest.mark.parametrize("variant", ["join01", "joinn"])
def test_join_aggregate_program(variant):
    b, p = datagen.gen_join_aggregate(5000, 20000, seed=4)
    db = relalg.Database().add_table("probe", p).add_table("build", b)
    plan = ("hashaggr", "hash",
            ("hashjoin", variant,
             ("select", ("le", "date", workloads.JOIN_DATE_CUTOFF), ("scan", "probe", ["fk", "date", "price"])),
             ["fk"], ("scan", "build", ["key", "segment"]), ["key"], ["segment"]),
            ["segment"], [("sum", "price")])
    text = V.program(plan)
    assert "bucket_lookup" in text and "bucket_insert" in text and "bucket_flush" in text
    dp, dr = both(db, plan)
    assert dp.njoins == 1 and dp.joins[0].variant == (abi.VH_JOIN01 if variant == "join01" else abi.VH_JOINN)
    assert dp.joins[0].npayload == 1
    want = oracle.sorted_rows(oracle.bruteforce(workloads.join_aggregate_pipeline(b, p).desc()))
    assert rows(dp) == rows(dr) == want


def test_star_program_four_joins_with_filtered_builds():
    f, dd, c, s, pt = datagen.gen_star(200_000, zipf_part=1.0, dims_scale=0.05, seed=5)
    db = relalg.Database()
    for name, t in (("fact", f), ("date", dd), ("customer", c), ("supplier", s), ("part", pt)):
        db.add_table(name, t)
    r = workloads.STAR_REGION
    fact = ("scan", "fact", ["f_date", "f_cust", "f_supp", "f_part", "f_revenue", "f_cost"])
    j1 = ("hashjoin", "joinn", fact, ["f_cust"],
          ("select", ("eq", "c_region", r), ("scan", "customer", ["c_key", "c_region", "c_nation"])),
          ["c_key"], ["c_nation"])
    j2 = ("hashjoin", "joinn", j1, ["f_supp"],
          ("select", ("eq", "s_region", r), ("scan", "supplier", ["s_key", "s_region", "s_nation"])),
          ["s_key"], ["s_nation"])
    j3 = ("hashjoin", "joinn", j2, ["f_date"], ("scan", "date", ["d_key", "d_year"]), ["d_key"], ["d_year"])
    j4 = ("hashjoin", "joinn", j3, ["f_part"], ("scan", "part", ["p_key"]), ["p_key"], [])
    plan = ("hashaggr", "hash", j4, ["c_nation", "s_nation", "d_year"],
            [("sum", ("-", "f_revenue", "f_cost")), ("count",)])
    dp, dr = both(db, plan)
    assert dp.njoins == 4 and dp.joins[0].build.npreds == 1 and dp.joins[3].npayload == 0
    want = oracle.sorted_rows(oracle.bruteforce(workloads.star_pipeline(f, dd, c, s, pt).desc()))
    assert rows(dp) == rows(dr) == want and len(want) <= 175


def test_global_aggregate_with_projection():
    t = datagen.gen_filter_groupby(10_000, seed=6)
    db = relalg.Database().add_table("l", t)
    plan = ("hashaggr", "global",
            ("project", [("disc_price", ("*", "price", ("-", 100, "disc")))],
             ("select", ("and", ("ge", "flag", 1), ("lt", "status", 1)),
              ("scan", "l", ["flag", "status", "price", "disc"]))),
            [], [("sum", "disc_price"), ("count",)])
    text = V.program(plan)
    assert "aggr_gsum" in text and "aggr_gcount" in text
    dp = db.lower_program(text)
    m = (t["flag"] >= 1) & (t["status"] < 1)
    want = [f"{int((t['price'][m].astype(object) * (100 - t['disc'][m])).sum())}|{int(m.sum())}"]
    assert dp.nkeys == 0 and dp.probe.npreds == 2
    assert rows(dp) == want


def test_multi_column_key_join_program():
    rng = np.random.default_rng(12)
    ps = {"ps_partkey": rng.integers(1, 200, 3000).astype(np.int32),
          "ps_suppkey": rng.integers(1, 40, 3000).astype(np.int32),
          "ps_cost": rng.integers(100, 1000, 3000).astype(np.int64),
          "ps_qty": rng.integers(1, 10, 3000).astype(np.int32)}
    li = {"l_partkey": rng.integers(1, 210, 20000).astype(np.int32),
          "l_suppkey": rng.integers(1, 42, 20000).astype(np.int32),
          "l_qty": rng.integers(1, 50, 20000).astype(np.int64)}
    db = relalg.Database().add_table("lineitem", li).add_table("partsupp", ps)
    plan = ("hashaggr", "hash",
            ("hashjoin", "joinn", ("scan", "lineitem", ["l_partkey", "l_suppkey", "l_qty"]),
             ["l_partkey", "l_suppkey"],
             ("scan", "partsupp", ["ps_partkey", "ps_suppkey", "ps_cost", "ps_qty"]),
             ["ps_partkey", "ps_suppkey"], ["ps_cost", "ps_qty"]),
            ["ps_qty"], [("sum", ("*", "l_qty", "ps_cost")), ("count",)])
    text = V.program(plan)
    assert "rehash" in text
    dp, dr = both(db, plan)
    assert dp.joins[0].nkeys == 2 and dp.joins[0].npayload == 2
    assert rows(dp) == rows(dr) and len(rows(dp)) == 9


def test_unsupported_program_shapes_are_refused():
    t = datagen.gen_filter_groupby(100, seed=1)
    db = relalg.Database().add_table("l", t)
    for text in ["(program)", "(program (pipeline (lolepop x (done))))", "(pipeline)",
                 "(program (basetable l l (col flag flag value)) (pipeline (lolepop s (bogus))))"]:
        with pytest.raises(runtime.VoilaHipError):
            db.lower_program(text)
