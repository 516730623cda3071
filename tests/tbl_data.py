"""Synthetic TPC-H-format text for the importer tests: a seeded lineitem.tbl
with the reference schema's sixteen '|'-terminated fields, and the Q1 / Q6
shaped pipelines (TPC-H's definitions, decimals scaled by 100 as the
importer stores them)."""
import datetime

import numpy as np

from vldb_voila_amd import tbl
from vldb_voila_amd.pipeline import Pipeline, col, table_from_arrays

FLAGS = "RAN"
STATUS = "OF"
EPOCH = datetime.date(1992, 1, 1)
SPAN_DAYS = 2526                       # 1992-01-01 .. 1998-12-01


def write_lineitem(path, rows, seed=1):
    rng = np.random.default_rng(seed)
    qty = rng.integers(1, 51, rows)
    price = rng.integers(90000, 10500000, rows)
    disc = rng.integers(0, 11, rows)
    tax = rng.integers(0, 9, rows)
    flag = rng.integers(0, 3, rows)
    status = rng.integers(0, 2, rows)
    ship = rng.integers(0, SPAN_DAYS, rows)
    lines = []
    for i in range(rows):
        d = EPOCH + datetime.timedelta(days=int(ship[i]))
        fields = [str(i // 4 + 1), str(i % 997 + 1), str(i % 101 + 1), str(i % 4 + 1),
                  f"{qty[i]}.00", f"{price[i] // 100}.{price[i] % 100:02d}", f"0.{disc[i]:02d}", f"0.{tax[i]:02d}",
                  FLAGS[flag[i]], STATUS[status[i]], d.isoformat(), d.isoformat(), d.isoformat(),
                  "DELIVER IN PERSON", "TRUCK", "a comment"]
        lines.append("|".join(fields) + "|\n")
    with open(path, "w") as f:
        f.writelines(lines)


Q1_COLS = ["l_quantity", "l_extendedprice", "l_discount", "l_tax", "l_returnflag", "l_linestatus", "l_shipdate"]
Q1_CUTOFF = tbl.julian_day(1998, 9, 2)          # date '1998-12-01' - interval '90' day
Q6_FROM, Q6_TO = tbl.julian_day(1994, 1, 1), tbl.julian_day(1995, 1, 1)


def q1(arrays, bounds, addrs=None):
    li = table_from_arrays("lineitem", arrays, addrs, bounds)
    p = Pipeline(li)
    p.select("l_shipdate", "<=", Q1_CUTOFF)
    p.groupby(["l_returnflag", "l_linestatus"], [
        ("sum", col("l_quantity")),
        ("sum", col("l_extendedprice")),
        ("sum", col("l_extendedprice") * (100 - col("l_discount"))),
        ("sum", col("l_extendedprice") * (100 - col("l_discount")) * (100 + col("l_tax"))),
        ("count",),
    ])
    return p


def q6(arrays, bounds, addrs=None):
    li = table_from_arrays("lineitem", arrays, addrs, bounds)
    p = Pipeline(li)
    p.select("l_shipdate", ">=", Q6_FROM)
    p.select("l_shipdate", "<", Q6_TO)
    p.select("l_discount", ">=", 5)
    p.select("l_discount", "<=", 7)
    p.select("l_quantity", "<", 2400)
    p.groupby([], [("sum", col("l_extendedprice") * col("l_discount"))])
    return p


def q1_from_text(path):
    """Q1 straight from the text with Python integers: sorted result rows."""
    groups = {}
    for line in open(path):
        f = line.split("|")
        q, p, d, t = (int(round(float(x) * 100)) for x in f[4:8])
        y, m, dd = (int(x) for x in f[10].split("-"))
        if tbl.julian_day(y, m, dd) > Q1_CUTOFF:
            continue
        key = (ord(f[8]), ord(f[9]))
        g = groups.setdefault(key, [0, 0, 0, 0, 0])
        g[0] += q
        g[1] += p
        g[2] += p * (100 - d)
        g[3] += p * (100 - d) * (100 + t)
        g[4] += 1
    return sorted("|".join(str(v) for v in (*k, *g)) for k, g in groups.items())


def q6_from_text(path):
    s, n = 0, 0
    for line in open(path):
        f = line.split("|")
        q, p, d = (int(round(float(x) * 100)) for x in f[4:7])
        y, m, dd = (int(x) for x in f[10].split("-"))
        j = tbl.julian_day(y, m, dd)
        if Q6_FROM <= j < Q6_TO and 5 <= d <= 7 and q < 2400:
            s += p * d
            n += 1
    return s, n


COLORS = ["almond", "antique", "aquamarine", "azure", "beige", "bisque", "black", "blanched", "blue", "blush",
          "brown", "burlywood", "chartreuse", "chiffon", "chocolate", "coral", "forest", "green", "honeydew",
          "khaki", "lavender", "lime", "navy", "olive", "orchid"]


def write_part(path, rows, seed=1):
    """part.tbl in the reference schema's nine fields: p_name five colour
    words (as dbgen draws them), p_mfgr / p_brand with TPC-H's patterns."""
    rng = np.random.default_rng(seed)
    lines = []
    for k in range(1, rows + 1):
        name = " ".join(COLORS[c] for c in rng.choice(len(COLORS), 5, replace=False))
        m = int(rng.integers(1, 6))
        fields = [str(k), name, f"Manufacturer#{m}", f"Brand#{m}{int(rng.integers(1, 6))}",
                  "STANDARD POLISHED TIN", str(int(rng.integers(1, 51))), "SM BOX",
                  f"{900 + k % 1000}.{k % 100:02d}", "regular deposits"]
        lines.append("|".join(fields) + "|\n")
    with open(path, "w") as f:
        f.writelines(lines)


def green_revenue_from_text(lineitem_path, part_path, word="green"):
    """sum(l_extendedprice * (100 - l_discount)) and count by l_returnflag
    over lineitems whose part's p_name contains `word`, from the text."""
    green = {int(f[0]) for f in (line.split("|") for line in open(part_path)) if word in f[1]}
    groups = {}
    for line in open(lineitem_path):
        f = line.split("|")
        if int(f[1]) not in green:
            continue
        p, d = (int(round(float(x) * 100)) for x in (f[5], f[6]))
        g = groups.setdefault(ord(f[8]), [0, 0])
        g[0] += p * (100 - d)
        g[1] += 1
    return sorted(f"{k}|{s}|{n}" for k, (s, n) in groups.items())


def green_revenue(li, part_keys, green, li_bounds, addrs=None, green_addr=None):
    """The same query as a pipeline: part filtered on a u8 predicate column
    (p_green == 1, the vec_contains result) joined with lineitem."""
    pt = table_from_arrays("part", {"p_partkey": part_keys, "p_green": green},
                           {"p_partkey": addrs["p_partkey"], "p_green": green_addr} if addrs else None,
                           {"p_green": (0, 1)})
    lt = table_from_arrays("lineitem", li, addrs["lineitem"] if addrs else None, li_bounds)
    p = Pipeline(lt)
    p.join(pt, "p_partkey", "l_partkey", build_preds=[("p_green", "==", 1)])
    p.groupby(["l_returnflag"], [("sum", col("l_extendedprice") * (100 - col("l_discount"))), ("count",)])
    return p


SEGMENTS = ["AUTOMOBILE", "BUILDING", "FURNITURE", "MACHINERY", "HOUSEHOLD"]
Q3_DATE = tbl.julian_day(1995, 3, 15)


def write_customer(path, rows, seed=1):
    """customer.tbl (eight fields): c_mktsegment one of TPC-H's five."""
    rng = np.random.default_rng(seed)
    lines = []
    for k in range(1, rows + 1):
        fields = [str(k), f"Customer#{k:09d}", "an address", str(int(rng.integers(0, 25))), "11-111-111-1111",
                  "100.00", SEGMENTS[int(rng.integers(0, 5))], "a comment"]
        lines.append("|".join(fields) + "|\n")
    with open(path, "w") as f:
        f.writelines(lines)


def write_orders(path, rows, customers, seed=1):
    """orders.tbl (nine fields) with orderkeys 1..rows, the keys
    write_lineitem's l_orderkey (row // 4 + 1) refers to."""
    rng = np.random.default_rng(seed)
    lines = []
    for k in range(1, rows + 1):
        d = EPOCH + datetime.timedelta(days=int(rng.integers(0, SPAN_DAYS)))
        fields = [str(k), str(int(rng.integers(1, customers + 1))), "O", "1000.00", d.isoformat(), "1-URGENT",
                  "Clerk#000000001", "0", "a comment"]
        lines.append("|".join(fields) + "|\n")
    with open(path, "w") as f:
        f.writelines(lines)


def q3_from_text(lineitem_path, orders_path, customer_path, segment="BUILDING"):
    """TPC-H Q3 from the text with Python integers: rows
    l_orderkey|o_orderdate|o_shippriority|revenue (revenue scaled by 10^4)."""
    cust = {int(f[0]) for f in (line.split("|") for line in open(customer_path)) if f[6] == segment}
    orders = {}
    for line in open(orders_path):
        f = line.split("|")
        y, m, d = (int(x) for x in f[4].split("-"))
        od = tbl.julian_day(y, m, d)
        if od < Q3_DATE and int(f[1]) in cust:
            orders[int(f[0])] = (od, int(f[7]))
    groups = {}
    for line in open(lineitem_path):
        f = line.split("|")
        y, m, d = (int(x) for x in f[10].split("-"))
        if tbl.julian_day(y, m, d) <= Q3_DATE or int(f[0]) not in orders:
            continue
        p, dc = (int(round(float(x) * 100)) for x in (f[5], f[6]))
        key = (int(f[0]), *orders[int(f[0])])
        groups[key] = groups.get(key, 0) + p * (100 - dc)
    return sorted(f"{k[0]}|{k[1]}|{k[2]}|{v}" for k, v in groups.items())


def q3(li, li_b, orders, o_b, customer, c_b, addrs=None):
    """TPC-H Q3 as one pipeline: lineitem (l_shipdate > date) joins orders
    (o_orderdate < date; payloads o_orderdate, o_shippriority, o_custkey),
    then customer on o_custkey (a u8 c_building predicate column, the
    varchar equality's result for c_mktsegment = 'BUILDING'); group by
    (l_orderkey, o_orderdate, o_shippriority): sum(l_extendedprice * (100 -
    l_discount)).  The reference plan joins customer into orders first
    (bench_tpch_rel.cpp:235-323); with unique keys on both build sides the
    rows are the same."""
    a = addrs or {}
    lt = table_from_arrays("lineitem", li, a.get("lineitem"), li_b)
    ot = table_from_arrays("orders", orders, a.get("orders"), o_b)
    ct = table_from_arrays("customer", customer, a.get("customer"), c_b)
    from vldb_voila_amd import abi
    p = Pipeline(lt)
    p.select("l_shipdate", ">", Q3_DATE)
    p.join(ot, "o_orderkey", "l_orderkey", payload=["o_orderdate", "o_shippriority", "o_custkey"],
           variant=abi.VH_JOIN01, build_preds=[("o_orderdate", "<", Q3_DATE)],
           names=["o_orderdate", "o_shippriority", "o_custkey"])
    p.join(ct, "c_custkey", "o_custkey", payload=[], variant=abi.VH_JOIN01, build_preds=[("c_building", "==", 1)],
           names=[])
    p.groupby(["l_orderkey", "o_orderdate", "o_shippriority"],
              [("sum", col("l_extendedprice") * (100 - col("l_discount")))])
    return p
