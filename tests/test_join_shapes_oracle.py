"""Join shapes beyond one key and one payload, on the CPU: the oracle's
vector(1024) restatement (hash, then rehash per further key column, as the
translator hashes key lists; the probe loop fans out at every JoinN) against
the brute force (sorted key tuples) for multi-column keys, several payload
columns, and JoinN fan-out before the last join; and the lowering's packing
limits (key and payload bit fields within 32 bits)."""
import ctypes as C

import numpy as np
import pytest

import oracle
from vldb_voila_amd import abi, runtime
from vldb_voila_amd.pipeline import Pipeline, col, table_from_arrays


def composite_pipeline(rng, n_build=3000, n_probe=20000, variant=abi.VH_JOINN):
    a = rng.integers(0, 40, n_build).astype(np.int32)
    b = rng.integers(-5, 30, n_build).astype(np.int32)
    build = {"a": a, "b": b, "p1": rng.integers(0, 9, n_build).astype(np.int32),
             "p2": rng.integers(0, 200, n_build).astype(np.int64),
             "p3": rng.integers(1000, 1003, n_build).astype(np.int16)}
    probe = {"x": rng.integers(0, 45, n_probe).astype(np.int32),
             "y": rng.integers(-8, 33, n_probe).astype(np.int32),
             "v": rng.integers(0, 10**6, n_probe).astype(np.int64)}
    r = table_from_arrays("r", build)
    s = table_from_arrays("s", probe)
    p = Pipeline(s).join(r, ["a", "b"], ["x", "y"], payload=["p1", "p2", "p3"], variant=variant,
                         names=["p1", "p2", "p3"])
    p.groupby(["p1", "p3"], [("sum", col("v") * col("p2")), ("count",)])
    return p.desc(), (build, probe)    # the arrays must outlive the descriptor


@pytest.mark.parametrize("seed", [1, 2])
def test_multi_column_keys_and_payloads_vector_flavor_equals_bruteforce(seed):
    d, keep = composite_pipeline(np.random.default_rng(seed))
    want = oracle.sorted_rows(oracle.bruteforce(d))
    assert len(want) > 5
    for vs in (1024, 16):
        got = oracle.sorted_rows(oracle.run_pipeline(d, threads=3, vector_size=vs)[0])
        assert got == want


def test_multi_column_key_bruteforce_by_hand():
    """Tiny case against a dictionary join written out here: (a, b) pairs
    repeat, and only exact pairs match (a alone or b alone does not)."""
    build = {"a": np.array([1, 1, 2, 2, 1], np.int32), "b": np.array([5, 6, 5, 6, 5], np.int32),
             "p": np.array([10, 20, 30, 40, 50], np.int32)}
    probe = {"x": np.array([1, 1, 2, 3, 1, 2], np.int32), "y": np.array([5, 6, 6, 5, 7, 5], np.int32),
             "v": np.array([1, 2, 3, 4, 5, 6], np.int64)}
    r, s = table_from_arrays("r", build), table_from_arrays("s", probe)
    d = Pipeline(s).join(r, ["a", "b"], ["x", "y"], payload=["p"], names=["p"]) \
        .groupby(["p"], [("sum", col("v")), ("count",)]).desc()
    # (1,5) -> p 10 and 50; (1,6) -> 20; (2,6) -> 40; (3,5), (1,7) none; (2,5) -> 30
    want = ["10|1|1", "20|2|1", "30|6|1", "40|3|1", "50|1|1"]
    assert oracle.sorted_rows(oracle.bruteforce(d)) == want
    assert oracle.sorted_rows(oracle.run_pipeline(d, threads=2)[0]) == want


def test_joinn_fan_out_before_the_last_join():
    """A JoinN with repeated keys first, then a second join: every match of
    the first join continues through the second."""
    rng = np.random.default_rng(7)
    d1 = {"k": rng.integers(0, 50, 400).astype(np.int32), "g": rng.integers(0, 6, 400).astype(np.int32)}
    d2 = {"k": np.arange(0, 100, dtype=np.int32), "h": rng.integers(0, 4, 100).astype(np.int32)}
    f = {"a": rng.integers(0, 60, 30000).astype(np.int32), "b": rng.integers(0, 110, 30000).astype(np.int32),
         "v": rng.integers(-1000, 1000, 30000).astype(np.int64)}
    p = Pipeline(table_from_arrays("f", f))
    p.join(table_from_arrays("d1", d1), "k", "a", payload=["g"], variant=abi.VH_JOINN, names=["g"])
    p.join(table_from_arrays("d2", d2), "k", "b", payload=["h"], variant=abi.VH_JOIN01, names=["h"])
    d = p.groupby(["g", "h"], [("sum", col("v")), ("count",)]).desc()
    want = oracle.sorted_rows(oracle.bruteforce(d))
    # closed form: count per (g, h) = sum over fact rows of matches
    cnt = {}
    by_k = {}
    for k, g in zip(d1["k"], d1["g"]):
        by_k.setdefault(int(k), []).append(int(g))
    for a, b in zip(f["a"], f["b"]):
        if b >= 100:
            continue
        for g in by_k.get(int(a), []):
            key = (g, int(d2["h"][b]))
            cnt[key] = cnt.get(key, 0) + 1
    assert {(int(r.split("|")[0]), int(r.split("|")[1])): int(r.split("|")[3]) for r in want} == cnt
    assert oracle.sorted_rows(oracle.run_pipeline(d, threads=4)[0]) == want


def _lower_rc(desc):
    L = runtime.lib()
    L.vh_generate_source.argtypes = [C.c_void_p, C.c_void_p, C.c_char_p, C.c_size_t, C.POINTER(C.c_size_t)]
    n = C.c_size_t()
    cfg = runtime.make_config()
    return L.vh_generate_source(C.byref(desc), C.byref(cfg), None, 0, C.byref(n)), L.vh_last_error()


def test_packing_limits():
    rng = np.random.default_rng(3)
    d, keep = composite_pipeline(rng)
    assert _lower_rc(d)[0] == 0
    # key columns of 20 + 20 bits: more than 32 bits together
    build = {"a": np.array([0, 1 << 19], np.int32), "b": np.array([0, 1 << 19], np.int32),
             "p": np.array([0, 1], np.int32)}
    probe = {"x": np.array([0], np.int32), "y": np.array([0], np.int32), "v": np.array([1], np.int64)}
    r, s = table_from_arrays("r", build), table_from_arrays("s", probe)
    wide = Pipeline(s).join(r, ["a", "b"], ["x", "y"], payload=["p"], names=["p"]) \
        .groupby(["p"], [("count",)]).desc()
    rc, msg = _lower_rc(wide)
    assert rc == 2 and b"key columns span more than 32 bits" in msg
    # payload columns of 17 + 17 bits
    build2 = {"k": np.array([0, 1], np.int32), "p": np.array([0, 1 << 16], np.int32),
              "q": np.array([0, 1 << 16], np.int32)}
    r2 = table_from_arrays("r", build2)
    s2 = table_from_arrays("s", {"x": np.array([0], np.int32), "v": np.array([1], np.int64)})
    wide2 = Pipeline(s2).join(r2, "k", "x", payload=["p", "q"], names=["p", "q"]) \
        .groupby(["p"], [("sum", col("q"))]).desc()
    rc, msg = _lower_rc(wide2)
    assert rc == 2 and b"payload columns span more than 32 bits" in msg
