"""The benchmark pipelines, as vh_pipeline_desc over given column addresses."""
from __future__ import annotations

from . import abi
from .datagen import SPAN_START
from .pipeline import Pipeline, col, table_from_arrays

FILTER_GROUPBY_CUTOFF = SPAN_START + 2435   # keeps 2436 of 2526 days (~96.4%)
JOIN_DATE_CUTOFF = SPAN_START + 1262        # keeps 1263 of 2526 days (50%)


def filter_groupby_pipeline(t: dict, addrs=None) -> Pipeline:
    """select date <= cutoff; group by (flag, status):
    sum(qty), sum(price), sum(price * (100 - disc)), count."""
    lineitem = table_from_arrays("lineitem", t, addrs)
    p = Pipeline(lineitem)
    p.select("date", "<=", FILTER_GROUPBY_CUTOFF)
    p.groupby(["flag", "status"], [
        ("sum", col("qty")),
        ("sum", col("price")),
        ("sum", col("price") * (100 - col("disc"))),
        ("count",),
    ])
    return p


def join_aggregate_pipeline(build: dict, probe: dict, build_addrs=None, probe_addrs=None,
                            variant=abi.VH_JOINN) -> Pipeline:
    """select date <= cutoff; join fk = key; group by segment: sum(price)."""
    r = table_from_arrays("build", build, build_addrs)
    s = table_from_arrays("probe", probe, probe_addrs)
    p = Pipeline(s)
    p.select("date", "<=", JOIN_DATE_CUTOFF)
    p.join(r, "key", "fk", payload=["segment"], variant=variant, names=["segment"])
    p.groupby(["segment"], [("sum", col("price"))])
    return p


STAR_REGION = 2


def star_pipeline(fact, date, customer, supplier, part, addrs=None, region=STAR_REGION) -> Pipeline:
    """Star join: fact x customer (region = r) x supplier (region = r) x date
    x part; group by (c_nation, s_nation, year): sum(revenue - cost).
    The filtered dimensions are joined first; part (no filter, no payload)
    last, so only surviving tuples probe its 1.4M-key table."""
    a = addrs or {}
    f = table_from_arrays("fact", fact, a.get("fact"))
    p = Pipeline(f)
    p.join(table_from_arrays("customer", customer, a.get("customer")), "c_key", "f_cust",
           payload=["c_nation"], build_preds=[("c_region", "==", region)], names=["c_nation"])
    p.join(table_from_arrays("supplier", supplier, a.get("supplier")), "s_key", "f_supp",
           payload=["s_nation"], build_preds=[("s_region", "==", region)], names=["s_nation"])
    p.join(table_from_arrays("date", date, a.get("date")), "d_key", "f_date",
           payload=["d_year"], names=["d_year"])
    p.join(table_from_arrays("part", part, a.get("part")), "p_key", "f_part", payload=[],
           names=[])
    p.groupby(["c_nation", "s_nation", "d_year"], [("sum", col("f_revenue") - col("f_cost")), ("count",)])
    return p


def batch_pipeline(params, build, probe, build_addrs=None, probe_addrs=None) -> Pipeline:
    """select a < selectivity * 1000; join fk = key (JoinN); group by g:
    sum(v), sum(w), count."""
    cut = int(round(params["selectivity"] * 1000))
    r = table_from_arrays("build", build, build_addrs)
    s = table_from_arrays("probe", probe, probe_addrs)
    p = Pipeline(s)
    p.select("a", "<", cut)
    p.join(r, "key", "fk", payload=["w"], variant=abi.VH_JOINN, names=["w"])
    p.groupby(["g"], [("sum", col("v")), ("sum", col("w")), ("count",)])
    return p
