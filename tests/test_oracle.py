"""Pins the CPU oracle: against the reference's own runtime (golden vectors
written by tests/golden/make_ref_vectors.py from oracle/_ref, and the live
oracle/_ref library when it is built), exact Python arithmetic, closed forms,
and an independent brute-force evaluation of whole pipelines."""
import ctypes as C
import json
import os

import numpy as np
import pytest

import oracle
from oracle import ref
from vldb_voila_amd import abi, datagen, workloads
from vldb_voila_amd.pipeline import Pipeline, col, table_from_arrays

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "ref_vectors.json")
M64 = (1 << 64) - 1


def py_hash(v):
    """voila_hash<T>::hash (runtime.hpp:413-424) in exact integers."""
    x = v & M64
    x ^= x >> 16
    x = (x * 0x45d9f3b) & M64
    x ^= x >> 16
    x = (x * 0x45d9f3b) & M64
    x ^= x >> 16
    return x


@pytest.fixture(scope="module")
def golden():
    with open(GOLDEN) as f:
        return json.load(f)


def test_hash_matches_reference_vectors(golden):
    L = oracle.lib()
    for v, h in golden["hash_i32"] + golden["hash_i64"]:
        assert L.vo_hash(v) == h
        assert py_hash(v) == h
    for v, h in golden["hash_u32"]:
        assert L.vo_hash(v) == h
    for h0, v, h in golden["rehash_i32"]:
        assert L.vo_rehash(h0, v) == h
        assert (py_hash(v) ^ ((h0 * 2654435761) & M64)) == h


def test_hash_zero_is_zero():
    # hash(0) == 0, so the all-zero group key hashes to 0 (partition 0); the
    # reference's debug-build ASSERT(hash != 0) in partition_data would fire.
    assert oracle.lib().vo_hash(0) == 0
    assert oracle.lib().vo_rehash(0, 0) == 0


def test_to_cstr_matches_reference_vectors(golden):
    L = oracle.lib()
    buf = C.create_string_buffer(64)
    for v, s in golden["to_cstr_i128"]:
        v = int(v)
        u = v & ((1 << 128) - 1)
        L.vo_to_cstr_i128(u & M64, u >> 64, buf)
        assert buf.value.decode() == s == str(v)
    for v, s in golden["to_cstr_i64"]:
        assert s == str(v)
    for v, s in golden["to_cstr_u64"]:
        assert s == v


def test_selection_matches_reference_vectors(golden):
    L = oracle.lib()
    for case in golden["select"]:
        pred = np.array(case["pred"], dtype=np.uint8)
        sel = None if case["sel"] is None else np.array(case["sel"], dtype=np.int32)
        res = np.zeros(1024, dtype=np.int32)
        sp = sel.ctypes.data if sel is not None else None
        k = L.vo_seltrue(sp, case["n"], res.ctypes.data, pred.ctypes.data)
        assert res[:k].tolist() == case["true"]
        k = L.vo_selfalse(sp, case["n"], res.ctypes.data, pred.ctypes.data)
        assert res[:k].tolist() == case["false"]
        # closed form: ascending, disjoint, covering the input selection
        base = list(range(case["n"])) if sel is None else case["sel"]
        assert sorted(case["true"] + case["false"]) == sorted(base)


def test_selunion_matches_reference_vectors(golden):
    L = oracle.lib()
    for case in golden["selunion"]:
        l = np.array(case["l"], dtype=np.int32)
        r = np.array(case["r"], dtype=np.int32)
        res = np.zeros(len(l) + len(r) + 1, dtype=np.int32)
        k = L.vo_selunion(res.ctypes.data, l.ctypes.data, len(l), r.ctypes.data, len(r))
        assert res[:k].tolist() == case["union"] == sorted(set(case["l"]) | set(case["r"]))


@pytest.mark.skipif(not ref.available(), reason="reference sources not present")
def test_oracle_against_live_reference_runtime():
    R = ref.lib()
    L = oracle.lib()
    rng = np.random.default_rng(7)
    for v in rng.integers(-2**31, 2**31, 2000):
        v = int(v)
        assert L.vo_hash(v) == R.ref_hash_i32(v)
        assert L.vo_hash(v & 0xFFFFFFFF) == R.ref_hash_u32(v & 0xFFFFFFFF)
    for _ in range(200):
        pred = (rng.random(1024) < rng.random()).astype(np.uint8)
        n = int(rng.integers(0, 1025))
        sel = np.sort(rng.choice(1024, size=n, replace=False)).astype(np.int32)
        a = np.zeros(1024, dtype=np.int32)
        b = np.zeros(1024, dtype=np.uint32)
        ka = L.vo_seltrue(sel.ctypes.data, n, a.ctypes.data, pred.ctypes.data)
        kb = R.ref_select_true(sel.ctypes.data, n, b.ctypes.data, pred.ctypes.data)
        assert ka == kb and (a[:ka] == b[:kb].astype(np.int32)).all()


@pytest.mark.skipif(not ref.available(), reason="reference sources not present")
def test_avx512_hash_differs_from_scalar_only_for_negative_keys():
    """The reference's AVX-512 hash path (runtime.cpp:768-799) shifts
    arithmetically; results never depend on hash values, only bucket choice."""
    R = ref.lib()
    vals = np.array([0, 1, 5, 1 << 20, 2**31 - 1, 123456], dtype=np.uint32)
    res = np.zeros(len(vals), dtype=np.uint64)
    R.ref_avx512_hash_u32(None, len(vals), res.ctypes.data, vals.ctypes.data)
    for v, h in zip(vals.tolist(), res.tolist()):
        assert h == py_hash(v)


@pytest.mark.skipif(not ref.available(), reason="reference sources not present")
def test_reference_bucket_lookup_next_check_semantics():
    """bucket_lookup = buckets[hash & mask]; bucket_next = *(row + next_off);
    check = *(u32*)(row + off) == key (runtime.cpp:578-723): the layout the
    oracle's chained tables follow."""
    R = ref.lib()
    rows = np.zeros((8, 4), dtype=np.uint64)        # row: key | pay | hash | next
    buckets = np.zeros(16, dtype=np.uint64)
    base = rows.ctypes.data
    for i in range(8):
        key = 100 + i
        h = py_hash(key)
        rows[i, 0] = key
        rows[i, 2] = h
        idx = h & 15
        rows[i, 3] = buckets[idx]
        buckets[idx] = base + i * 32
    hashes = np.array([py_hash(100 + i) for i in range(8)], dtype=np.uint64)
    res = np.zeros(8, dtype=np.uint64)
    R.ref_bucket_lookup(None, 8, res.ctypes.data, buckets.ctypes.data, hashes.ctypes.data, 15)
    assert (res == buckets[hashes & np.uint64(15)]).all()
    keys = np.array([100 + i for i in range(8)], dtype=np.uint32)
    chk = np.zeros(8, dtype=np.uint8)
    # walk each chain until the key is found
    cur = res.copy()
    found = np.zeros(8, dtype=bool)
    for _ in range(8):
        live = np.nonzero(cur)[0].astype(np.int32)
        if not len(live):
            break
        chk[:] = 0
        R.ref_check_u32(live.ctypes.data, len(live), chk.ctypes.data, cur.ctypes.data,
                        keys.ctypes.data, 0)
        found |= (chk == 1) & (cur != 0)
        nxt = np.zeros(8, dtype=np.uint64)
        R.ref_bucket_next(live.ctypes.data, len(live), nxt.ctypes.data, cur.ctypes.data, 24)
        cur = np.where(cur != 0, nxt, 0).astype(np.uint64)
    assert found.all()


def test_calc_num_buckets_closed_form():
    L = oracle.lib()
    for count in list(range(0, 5000, 37)) + [1_400_000, 1_500_000, 6_000_000]:
        n = max(count, 2048, 2048)
        exp = n.bit_length()
        if (1 << exp) < n / 0.7:
            exp += 1
        assert L.vo_calc_num_buckets(count, 1024, 2048) == 1 << exp
    assert L.vo_calc_num_buckets(0, 1024, 2048) == 4096
    assert L.vo_calc_num_buckets(1_500_000, 1024, 2048) == 4_194_304
    assert L.vo_calc_num_buckets(1_400_000, 1024, 2048) == 2_097_152


def test_type_from_minmax_boundaries():
    L = oracle.lib()
    assert L.vo_type_from_minmax(0, 255) == abi.VH_U8
    assert L.vo_type_from_minmax(-1, 127) == abi.VH_I8
    assert L.vo_type_from_minmax(0, 256) == abi.VH_U16
    assert L.vo_type_from_minmax(-1, 2**31 - 1) == abi.VH_I32
    assert L.vo_type_from_minmax(0, 2**32 - 1) == abi.VH_U32
    assert L.vo_type_from_minmax(0, 2**63) == abi.VH_U64
    # the limits are compared as doubles: (double)INT64_MAX rounds to 2^63
    assert L.vo_type_from_minmax(-1, 2**63) == abi.VH_I64
    assert L.vo_type_from_minmax(-1, 2**64) == abi.VH_I128
    assert L.vo_type_from_minmax(0, 2**65) == abi.VH_I128


def test_filter_groupby_typing():
    """price*(100-disc) is u32 and its sum i128 (typing_pass.cpp:164-176, :237-266)."""
    t = datagen.gen_filter_groupby(5000)
    d = workloads.filter_groupby_pipeline(t).desc()
    nodes = (C.c_int32 * abi.VH_MAX_EXPR)()
    ag = (C.c_int32 * abi.VH_MAX_AGGRS)()
    fa = (C.c_int32 * abi.VH_MAX_AGGRS)()
    assert oracle.lib().vo_typing_of(C.byref(d), nodes, ag, fa) == 0
    # nodes: qty, price, price(shared), 100, disc, 100-disc, price*(100-disc)
    types = list(nodes[:d.nexpr])
    root = d.aggrs[2].expr
    assert types[root] == abi.VH_U32
    assert ag[0] == abi.VH_U64 and ag[1] == abi.VH_U64 and ag[2] == abi.VH_I128
    assert ag[3] == abi.VH_U64


def _tiny_known_answer():
    probe = {
        "k": np.array([1, 2, 1, 3, 2, 1, 0, 3], dtype=np.int32),
        "v": np.array([10, -20, 30, 40, -50, 60, 70, 80], dtype=np.int64),
        "f": np.array([1, 1, 1, 0, 1, 1, 1, 1], dtype=np.int32),
    }
    p = Pipeline(table_from_arrays("t", probe))
    p.select("f", "==", 1)
    p.groupby(["k"], [("sum", col("v")), ("count",), ("sum", col("v") * 2 - 5)])
    return p


def test_known_answer_tiny():
    p = _tiny_known_answer()
    d = p.desc()
    want = sorted(["0|70|1|135", "1|100|3|185", "2|-70|2|-150", "3|80|1|155"])
    for threads in (1, 2, 3):
        got, _ = oracle.run_pipeline(d, threads=threads)
        assert oracle.sorted_rows(got) == want
    assert oracle.sorted_rows(oracle.bruteforce(d)) == want


def test_bucket_insert_conflicts_small_vectors():
    """Many groups per vector: bucket_insert conflicts and retries, table
    index rebuilds (runtime_struct.hpp:458-485) with vector size 16."""
    rng = np.random.default_rng(3)
    n = 20000
    probe = {"k1": rng.integers(0, 300, n).astype(np.int32),
             "k2": rng.integers(-3, 4, n).astype(np.int32),
             "v": rng.integers(-10**12, 10**12, n).astype(np.int64)}
    p = Pipeline(table_from_arrays("t", probe))
    p.groupby(["k1", "k2"], [("sum", col("v")), ("count",)])
    d = p.desc()
    want = oracle.sorted_rows(oracle.bruteforce(d))
    assert len(want) == len(set(zip(probe["k1"].tolist(), probe["k2"].tolist())))
    for vs, ms, th in ((16, 64, 1), (1024, 16384, 4), (7, 49, 3)):
        got, _ = oracle.run_pipeline(d, threads=th, vector_size=vs, morsel_size=ms)
        assert oracle.sorted_rows(got) == want


@pytest.mark.parametrize("threads", [1, 4])
def test_filter_groupby_vector_flavor_matches_bruteforce(threads):
    t = datagen.gen_filter_groupby(300_000, seed=5)
    d = workloads.filter_groupby_pipeline(t).desc()
    got, _ = oracle.run_pipeline(d, threads=threads)
    want = oracle.bruteforce(d)
    assert oracle.sorted_rows(got) == oracle.sorted_rows(want)
    # closed form in numpy for the totals
    m = t["date"] <= workloads.FILTER_GROUPBY_CUTOFF
    rows = [r.split("|") for r in oracle.sorted_rows(want)]
    assert sum(int(r[5]) for r in rows) == int(m.sum())
    assert sum(int(r[2]) for r in rows) == int(t["qty"][m].sum())
    exact = sum(int(a) * (100 - int(b)) for a, b in zip(t["price"][m][:1000], t["disc"][m][:1000]))
    assert exact > 0


@pytest.mark.parametrize("variant", [abi.VH_JOIN01, abi.VH_JOINN])
def test_join_aggregate_vector_flavor_matches_bruteforce(variant):
    b, pr = datagen.gen_join_aggregate(40_000, 160_000, seed=9)
    d = workloads.join_aggregate_pipeline(b, pr, variant=variant).desc()
    got, _ = oracle.run_pipeline(d, threads=3)
    want = oracle.bruteforce(d)
    assert oracle.sorted_rows(got) == oracle.sorted_rows(want)
    # closed form: every probe key matches exactly one build row
    m = pr["date"] <= workloads.JOIN_DATE_CUTOFF
    seg_of = np.zeros(40_001, dtype=np.int64)
    seg_of[b["key"]] = b["segment"]
    sums = np.bincount(seg_of[pr["fk"][m]], weights=None, minlength=5)
    tot = {}
    for s, v in zip(seg_of[pr["fk"][m]].tolist(), pr["price"][m].tolist()):
        tot[s] = tot.get(s, 0) + v
    assert sorted(f"{s}|{v}" for s, v in tot.items()) == oracle.sorted_rows(want)
    assert sums.sum() == m.sum()


def test_joinn_duplicates_multiply_matches():
    build = {"key": np.array([5, 5, 6, 7, 7, 7], dtype=np.int32),
             "pay": np.array([1, 2, 3, 4, 5, 6], dtype=np.int32)}
    probe = {"fk": np.array([5, 6, 7, 8, 5], dtype=np.int32),
             "v": np.array([100, 200, 300, 400, 500], dtype=np.int64)}
    p = Pipeline(table_from_arrays("s", probe))
    p.join(table_from_arrays("r", build), "key", "fk", payload=["pay"], variant=abi.VH_JOINN,
           names=["pay"])
    p.groupby(["pay"], [("sum", col("v")), ("count",)])
    d = p.desc()
    want = sorted(["1|600|2", "2|600|2", "3|200|1", "4|300|1", "5|300|1", "6|300|1"])
    got, _ = oracle.run_pipeline(d, threads=2)
    assert oracle.sorted_rows(got) == want
    assert oracle.sorted_rows(oracle.bruteforce(d)) == want


def test_empty_and_fully_filtered_inputs():
    probe = {"k": np.array([], dtype=np.int32), "v": np.array([], dtype=np.int64)}
    p = Pipeline(table_from_arrays("t", probe))
    p.groupby(["k"], [("sum", col("v"))])
    got, _ = oracle.run_pipeline(p.desc(), threads=2)
    assert got == "" and oracle.bruteforce(p.desc()) == ""
    t = datagen.gen_filter_groupby(5000)
    p = workloads.filter_groupby_pipeline(t)
    p.preds[0] = (p.probe.index("date"), abi.VH_LT, 0)
    assert oracle.run_pipeline(p.desc())[0] == "" and oracle.bruteforce(p.desc()) == ""


def test_global_aggregate():
    t = datagen.gen_filter_groupby(50_000)
    p = Pipeline(table_from_arrays("t", t))
    p.select("flag", "==", 1)
    p.groupby([], [("sum", col("price") * col("qty")), ("count",)])
    d = p.desc()
    m = t["flag"] == 1
    want = [f"{int((t['price'][m].astype(object) * t['qty'][m]).sum())}|{int(m.sum())}"]
    assert oracle.sorted_rows(oracle.run_pipeline(d, threads=3)[0]) == want
    assert oracle.sorted_rows(oracle.bruteforce(d)) == want


def test_oracle_detects_a_dropped_term():
    """A plausible kernel mistake (dropping the (100 - disc) factor) changes
    the oracle's answer, so parity tests would catch it."""
    t = datagen.gen_filter_groupby(20_000)
    good = workloads.filter_groupby_pipeline(t)
    bad = workloads.filter_groupby_pipeline(t)
    bad.aggrs[2] = ("sum", col("price") * 100)
    assert oracle.sorted_rows(oracle.bruteforce(good.desc())) != \
        oracle.sorted_rows(oracle.bruteforce(bad.desc()))


def test_desc_layout_matches_c():
    assert oracle.lib().vo_sizeof_pipeline_desc() == C.sizeof(abi.vh_pipeline_desc)


def test_vector_flavor_with_colliding_hash_keys():
    """Build keys whose voila_hash shares its low 16 bits: every key lands in
    one bucket chain of the vector flavor's table (and the insert-conflict
    rule retries within it); its rows equal the brute force's."""
    m64 = np.uint64(0x45d9f3b)
    keys = [0]
    start = 1
    while len(keys) < 600:
        o = np.arange(start, start + (1 << 22), dtype=np.uint64)
        h = o ^ (o >> np.uint64(16))
        h = h * m64
        h = h ^ (h >> np.uint64(16))
        h = h * m64
        h = h ^ (h >> np.uint64(16))
        keys.extend(o[(h & np.uint64(0xFFFF)) == 0].tolist())
        start += 1 << 22
    keys = np.array(keys[:600], dtype=np.int64)
    assert all(oracle.lib().vo_hash(int(k)) & 0xFFFF == 0 for k in keys[:20])
    rng = np.random.default_rng(9)
    build = {"k": keys[:350], "p": rng.integers(0, 12, 350).astype(np.int32)}
    probe = {"fk": keys[rng.integers(0, 600, 20_000)], "v": rng.integers(-10**12, 10**12, 20_000)}
    for variant in (abi.VH_JOINN, abi.VH_JOIN01):
        p = Pipeline(table_from_arrays("s", probe))
        p.join(table_from_arrays("r", build), "k", "fk", payload=["p"], variant=variant, names=["p"])
        p.groupby(["p"], [("sum", col("v")), ("count",)])
        want = oracle.sorted_rows(oracle.bruteforce(p.desc()))
        for threads in (1, 4):
            got = oracle.sorted_rows(oracle.run_pipeline(p.desc(), threads=threads)[0])
            assert got == want, (variant, threads)
    # groups keyed by colliding values too: every group in one chain
    t = {"g": keys[rng.integers(0, 600, 30_000)], "v": rng.integers(0, 1000, 30_000)}
    p = Pipeline(table_from_arrays("t", t))
    p.groupby(["g"], [("sum", col("v")), ("count",)])
    want = oracle.sorted_rows(oracle.bruteforce(p.desc()))
    assert len(want) == 600
    assert oracle.sorted_rows(oracle.run_pipeline(p.desc(), threads=4)[0]) == want
