"""GPU parity on imported `.tbl` data: TPC-H Q1 and Q6 shaped pipelines over
lineitem columns loaded through the importer (decimals scaled by 100, dates
as Julian days, flags as bytes, the importer's min/max as metadata),
bit-exact against the oracle on the same arrays."""
import pytest

import oracle
from vldb_voila_amd import abi, runtime, tbl

from tbl_data import Q1_COLS, q1, q6, write_lineitem

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lineitem(tmp_path_factory):
    print("tbl: write and import a 200k-row lineitem.tbl", flush=True)
    d = tmp_path_factory.mktemp("tpch")
    write_lineitem(d / "lineitem.tbl", 200_000, seed=11)
    return tbl.load_table(str(d), "lineitem", Q1_COLS)


@pytest.mark.parametrize("kernel", [abi.VH_KERNEL_GENERIC, abi.VH_KERNEL_JIT, abi.VH_KERNEL_JIT_LAZY])
@pytest.mark.parametrize("query", ["q1", "q6"])
def test_tpch_shapes_on_imported_columns(lineitem, query, kernel):
    print(f"tbl: {query} kernel={kernel}", flush=True)
    a, b = lineitem
    make = q1 if query == "q1" else q6
    want = oracle.sorted_rows(oracle.bruteforce(make(a, b).desc()))
    dev = runtime.upload(a)
    got = oracle.sorted_rows(runtime.run_pipeline(make(a, b, runtime.addrs(dev)), runtime.make_config(kernel=kernel)))
    assert got == want
    if query == "q1":
        assert len(want) == 6


def test_relalg_tpch_plans_on_the_gpu(lineitem):
    """Q1 and Q6 as relalg plans over the imported catalog, run on the GPU."""
    print("tbl: relalg Q1 / Q6", flush=True)
    from vldb_voila_amd import relalg
    from tbl_data import Q1_CUTOFF, Q6_FROM, Q6_TO
    a, b = lineitem
    dev = runtime.upload(a)
    db = relalg.Database().add_table("lineitem", a, runtime.addrs(dev), bounds=b)
    host = relalg.Database().add_table("lineitem", a, bounds=b)
    for plan in (relalg.tpch_q1_plan(Q1_CUTOFF), relalg.tpch_q6_plan(Q6_FROM, Q6_TO)):
        want = oracle.sorted_rows(oracle.bruteforce(host.lower(plan)))
        q = db.query(plan)
        q.run()
        assert oracle.sorted_rows(q.result()) == want
        q.free()
