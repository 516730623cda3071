"""Join shapes beyond one key and one payload on the GPU, bit-exact against
the oracle: several payload columns packed into one entry word, keys of two
and three columns, JoinN fan-out before the last join (the generic kernel
fans out; a query on a generated kernel switches to it), and TPC-H Q3 as one
pipeline over imported .tbl columns, with c_mktsegment = 'BUILDING' computed
on the GPU by the varchar equality."""
import os

import numpy as np
import pytest

import oracle
from vldb_voila_amd import abi, runtime, tbl, varchar
from vldb_voila_amd.pipeline import Pipeline, col, table_from_arrays

import tbl_data as T

pytestmark = pytest.mark.gpu

KERNELS = [abi.VH_KERNEL_AUTO, abi.VH_KERNEL_GENERIC, abi.VH_KERNEL_JIT_LAZY]


@pytest.fixture(scope="module", autouse=True)
def device():
    print("gpu join shapes: init device 0", flush=True)
    runtime.init(0)


def composite(rng, n_build, n_probe, nkeys, variant=abi.VH_JOINN):
    build = {"a": rng.integers(0, 40, n_build).astype(np.int32),
             "b": rng.integers(-5, 30, n_build).astype(np.int32),
             "c": rng.integers(100, 104, n_build).astype(np.int16),
             "p1": rng.integers(0, 9, n_build).astype(np.int32),
             "p2": rng.integers(0, 200, n_build).astype(np.int64),
             "p3": rng.integers(1000, 1003, n_build).astype(np.int16)}
    probe = {"x": rng.integers(0, 45, n_probe).astype(np.int32),
             "y": rng.integers(-8, 33, n_probe).astype(np.int32),
             "z": rng.integers(99, 105, n_probe).astype(np.int16),
             "v": rng.integers(0, 10**6, n_probe).astype(np.int64)}
    bk, pk = ["a", "b", "c"][:nkeys], ["x", "y", "z"][:nkeys]

    def make(x=None, y=None):
        r = table_from_arrays("r", build, x)
        s = table_from_arrays("s", probe, y)
        p = Pipeline(s).join(r, bk, pk, payload=["p1", "p2", "p3"], variant=variant, names=["p1", "p2", "p3"])
        return p.groupby(["p1", "p3"], [("sum", col("v") * col("p2")), ("sum", col("p2")), ("count",)])
    return build, probe, make


@pytest.mark.parametrize("nkeys", [2, 3])
@pytest.mark.parametrize("layout", [abi.VH_LAYOUT_AUTO, abi.VH_LAYOUT_HASH])
def test_multi_column_keys_and_payloads(nkeys, layout):
    print(f"join shapes: {nkeys} key columns, layout {layout}", flush=True)
    rng = np.random.default_rng(10 + nkeys)
    build, probe, make = composite(rng, 40_000, 300_000, nkeys)
    want = oracle.sorted_rows(oracle.bruteforce(make().desc()))
    assert len(want) > 10
    devs = [runtime.upload(build), runtime.upload(probe)]
    p = make(*[runtime.addrs(d) for d in devs])
    for kernel in KERNELS:
        q = runtime.Query(p.desc(), runtime.make_config(kernel=kernel, join_layout=layout))
        if layout == abi.VH_LAYOUT_HASH:
            assert q.join(0)[0] == 8
        for rep in range(2):
            q.run()
            assert oracle.sorted_rows(q.result()) == want, (kernel, rep)


@pytest.mark.parametrize("kernel", KERNELS)
def test_joinn_fan_out_before_the_last_join(kernel):
    """A JoinN with repeated keys, then a Join01: every match of the first
    join continues through the second.  Generated kernels take first matches
    there and report it; the query then runs again on the generic kernel."""
    print(f"join shapes: JoinN fan-out before the last join, kernel {kernel}", flush=True)
    rng = np.random.default_rng(7)
    d1 = {"k": rng.integers(0, 5000, 40_000).astype(np.int32), "g": rng.integers(0, 6, 40_000).astype(np.int32)}
    d2 = {"k": np.arange(0, 10_000, dtype=np.int32), "h": rng.integers(0, 4, 10_000).astype(np.int32)}
    f = {"a": rng.integers(0, 6000, 500_000).astype(np.int32),
         "b": rng.integers(0, 11_000, 500_000).astype(np.int32),
         "v": rng.integers(-1000, 1000, 500_000).astype(np.int64)}

    def make(x=None, y=None, z=None):
        p = Pipeline(table_from_arrays("f", f, z))
        p.join(table_from_arrays("d1", d1, x), "k", "a", payload=["g"], variant=abi.VH_JOINN, names=["g"])
        p.join(table_from_arrays("d2", d2, y), "k", "b", payload=["h"], variant=abi.VH_JOIN01, names=["h"])
        return p.groupby(["g", "h"], [("sum", col("v")), ("count",)])
    want = oracle.sorted_rows(oracle.bruteforce(make().desc()))
    devs = [runtime.upload(t) for t in (d1, d2, f)]
    p = make(*[runtime.addrs(d) for d in devs])
    q = runtime.Query(p.desc(), runtime.make_config(kernel=kernel))
    for rep in range(2):
        q.run()
        assert oracle.sorted_rows(q.result()) == want, rep


@pytest.fixture(scope="module")
def tpch(tmp_path_factory):
    print("tbl: write and import lineitem / orders / customer", flush=True)
    d = tmp_path_factory.mktemp("q3")
    T.write_lineitem(d / "lineitem.tbl", 400_000, seed=3)
    T.write_orders(d / "orders.tbl", 100_000, 10_000, seed=4)
    T.write_customer(d / "customer.tbl", 10_000, seed=5)
    li = tbl.load_table(str(d), "lineitem", ["l_orderkey", "l_extendedprice", "l_discount", "l_shipdate"])
    od = tbl.load_table(str(d), "orders", ["o_orderkey", "o_custkey", "o_orderdate", "o_shippriority"])
    cu = tbl.load_table(str(d), "customer", ["c_custkey", "c_mktsegment"])
    text = T.q3_from_text(str(d / "lineitem.tbl"), str(d / "orders.tbl"), str(d / "customer.tbl"))
    return li, od, cu, text


@pytest.mark.parametrize("kernel", KERNELS)
def test_tpch_q3_over_imported_tables(tpch, kernel):
    print(f"tbl: Q3 kernel={kernel}", flush=True)
    (li, lb), (od, ob), (cu, cb), text = tpch
    seg = cu["c_mktsegment"]
    dseg = runtime.DeviceBuffer.from_numpy(np.ascontiguousarray(seg))
    flag = varchar.compare("==", dseg.ptr, len(seg), seg.shape[1] - 1, "BUILDING")
    building = flag.to_numpy(np.uint8, len(seg))
    # the GPU's equality agrees with the bytes
    word = np.frombuffer(b"BUILDING", np.uint8)
    assert np.array_equal(building, ((seg[:, 0] == 8) & np.all(seg[:, 1:9] == word, axis=1)).astype(np.uint8))
    cust = {"c_custkey": cu["c_custkey"], "c_building": building}
    cbb = {"c_custkey": cb["c_custkey"], "c_building": (0, 1)}
    want = oracle.sorted_rows(oracle.bruteforce(T.q3(li, lb, od, ob, cust, cbb).desc()))
    assert want == text and len(want) > 100
    dl, do = runtime.upload(li), runtime.upload(od)
    dc = runtime.upload({"c_custkey": cu["c_custkey"]})
    addrs = {"lineitem": runtime.addrs(dl), "orders": runtime.addrs(do),
             "customer": {"c_custkey": dc["c_custkey"].ptr, "c_building": flag.ptr}}
    p = T.q3(li, lb, od, ob, cust, cbb, addrs)
    q = runtime.Query(p.desc(), runtime.make_config(kernel=kernel))
    q.run()
    assert oracle.sorted_rows(q.result()) == want
