"""Statement-level parity: each vec_* kernel against the oracle's scalar
restatement or exact numpy, and the reference's group-by and join loops run
statement by statement on the GPU against the oracle."""
import ctypes as C

import numpy as np
import pytest

import oracle
from vldb_voila_amd import abi, datagen, runtime, statements, workloads
from vldb_voila_amd.runtime import DeviceBuffer

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def device():
    print("gpu statements: init device 0", flush=True)
    runtime.init(0)


def dev(a):
    return DeviceBuffer.from_numpy(np.ascontiguousarray(a))


@pytest.mark.parametrize("n,density,with_sel", [(0, 0.5, False), (1, 1.0, False), (4095, 0.3, False),
                                                (1_000_003, 0.5, False), (200_000, 0.9, True),
                                                (5000, 0.0, True)])
def test_seltrue_selfalse(n, density, with_sel):
    print(f"seltrue n={n}", flush=True)
    L = statements._bind()
    rng = np.random.default_rng(n)
    dom = max(n * 2, 1)
    pred = (rng.random(dom) < density).astype(np.uint8)
    sel = np.sort(rng.choice(dom, size=n, replace=False)).astype(np.int32) if with_sel else None
    dp, ds = dev(pred), (dev(sel) if sel is not None else None)
    res = DeviceBuffer(max(n, 1) * 4)
    cnt = DeviceBuffer(16)
    base = sel if sel is not None else np.arange(n, dtype=np.int32)
    for want, fn in ((1, L.vh_vec_seltrue), (0, L.vh_vec_selfalse)):
        runtime.check(fn(ds.ptr if ds else None, n, res.ptr, dp.ptr, cnt.ptr, None))
        k = int(cnt.to_numpy(np.int32, 1)[0])
        got = res.to_numpy(np.int32, k)
        exp = base[(pred[base] != 0) == bool(want)]
        assert np.array_equal(got, exp)
        # the oracle agrees on a 1024-position slice
        if n >= 1024 and sel is None:
            o = np.zeros(1024, dtype=np.int32)
            ok = (oracle.lib().vo_seltrue if want else oracle.lib().vo_selfalse)(
                None, 1024, o.ctypes.data, pred.ctypes.data)
            assert np.array_equal(o[:ok], exp[exp < 1024])


def test_selections_on_two_streams_do_not_share_scratch():
    """Two different selections enqueued back to back on two streams and
    synchronised only at the end: each keeps its own tile-count scratch, and
    both agree with the oracle's vec_seltrue / vec_selfalse."""
    print("seltrue on two streams", flush=True)
    L = statements._bind()
    rng = np.random.default_rng(7)
    n1, n2 = 3_000_017, 2_100_001
    p1 = (rng.random(n1) < 0.3).astype(np.uint8)
    p2 = (rng.random(n2) < 0.7).astype(np.uint8)
    d1, d2 = dev(p1), dev(p2)
    r1, r2 = DeviceBuffer(n1 * 4), DeviceBuffer(n2 * 4)
    c1, c2 = DeviceBuffer(16), DeviceBuffer(16)
    s1, s2 = runtime.Stream(), runtime.Stream()
    for _ in range(3):
        runtime.check(L.vh_vec_seltrue(None, n1, r1.ptr, d1.ptr, c1.ptr, s1.ptr))
        runtime.check(L.vh_vec_selfalse(None, n2, r2.ptr, d2.ptr, c2.ptr, s2.ptr))
    s1.synchronize()
    s2.synchronize()
    for n, pred, res, cnt, want in ((n1, p1, r1, c1, 1), (n2, p2, r2, c2, 0)):
        k = int(cnt.to_numpy(np.int32, 1)[0])
        got = res.to_numpy(np.int32, k)
        exp = np.zeros(n, dtype=np.int32)
        fn = oracle.lib().vo_seltrue if want else oracle.lib().vo_selfalse
        m = fn(None, n, exp.ctypes.data, pred.ctypes.data)
        assert k == m
        assert np.array_equal(got, exp[:m])


def test_hash_rehash_against_oracle():
    print("hash", flush=True)
    L = statements._bind()
    rng = np.random.default_rng(1)
    n = 100_000
    for dt, vt in ((np.int32, abi.VH_I32), (np.uint32, abi.VH_U32), (np.int64, abi.VH_I64)):
        info = np.iinfo(dt)
        v = rng.integers(info.min, info.max, n, dtype=dt)
        dv = dev(v)
        h = DeviceBuffer(n * 8)
        runtime.check(L.vh_vec_hash(vt, None, n, h.ptr, dv.ptr, None))
        got = h.to_numpy(np.uint64, n)
        for i in range(0, n, 997):
            assert int(got[i]) == oracle.lib().vo_hash(int(v[i]))
        runtime.check(L.vh_vec_rehash(vt, None, n, h.ptr, h.ptr, dv.ptr, None))
        got2 = h.to_numpy(np.uint64, n)
        for i in range(0, n, 997):
            assert int(got2[i]) == oracle.lib().vo_rehash(int(got[i]), int(v[i]))


def test_cmp_and_arith_casts():
    print("cmp/arith", flush=True)
    L = statements._bind()
    rng = np.random.default_rng(2)
    n = 50_000
    a = rng.integers(-1000, 1000, n).astype(np.int32)
    b = rng.integers(-1000, 1000, n).astype(np.int32)
    da, db = dev(a), dev(b)
    res = DeviceBuffer(n * 8)
    for cmp, fn in ((abi.VH_LT, np.less), (abi.VH_LE, np.less_equal), (abi.VH_GE, np.greater_equal),
                    (abi.VH_NE, np.not_equal)):
        runtime.check(L.vh_vec_cmp(cmp, abi.VH_I32, None, n, res.ptr, da.ptr, db.ptr, 0, None))
        assert np.array_equal(res.to_numpy(np.uint8, n), fn(a, b).astype(np.uint8))
        runtime.check(L.vh_vec_cmp(cmp, abi.VH_I32, None, n, res.ptr, da.ptr, None, 17, None))
        assert np.array_equal(res.to_numpy(np.uint8, n), fn(a, 17).astype(np.uint8))
    # operands are cast to the result type first: i16 result wraps
    for op, fn in ((abi.VH_E_ADD, np.add), (abi.VH_E_SUB, np.subtract), (abi.VH_E_MUL, np.multiply)):
        runtime.check(L.vh_vec_arith(op, abi.VH_I16, None, n, res.ptr, abi.VH_I32, da.ptr, abi.VH_I32, db.ptr,
                                     0, None))
        exp = fn(a.astype(np.int16), b.astype(np.int16))
        assert np.array_equal(res.to_numpy(np.int16, n), exp)
        runtime.check(L.vh_vec_arith(op, abi.VH_I64, None, n, res.ptr, abi.VH_I32, da.ptr, abi.VH_I32, db.ptr,
                                     0, None))
        assert np.array_equal(res.to_numpy(np.int64, n), fn(a.astype(np.int64), b.astype(np.int64)))


def test_groupby_statement_loop_matches_oracle():
    print("group-by statement loop", flush=True)
    t = datagen.gen_filter_groupby(200_000, seed=4)
    dt = runtime.upload(t)
    tbl, rows = statements.groupby_statements(
        [(abi.VH_I32, dt["flag"].ptr), (abi.VH_I32, dt["status"].ptr)],
        [(abi.VH_I64, dt["qty"].ptr), (abi.VH_I64, dt["price"].ptr), (abi.VH_I64, None)], 200_000)
    assert rows == 6
    k0 = tbl.column(0, rows)
    k1 = tbl.column(1, rows)
    s0, s1, c = tbl.column(2, rows), tbl.column(3, rows), tbl.column(4, rows)
    got = sorted(f"{a}|{b}|{x}|{y}|{z}" for a, b, x, y, z in zip(k0, k1, s0, s1, c))
    from vldb_voila_amd.pipeline import Pipeline, col, table_from_arrays
    p = Pipeline(table_from_arrays("t", t))
    p.groupby(["flag", "status"], [("sum", col("qty")), ("sum", col("price")), ("count",)])
    assert got == oracle.sorted_rows(oracle.bruteforce(p.desc()))


def test_groupby_statement_loop_many_groups_conflicts():
    print("group-by statement loop, many groups", flush=True)
    rng = np.random.default_rng(8)
    n = 100_000
    k = rng.integers(0, 20_000, n).astype(np.int32)
    v = rng.integers(-10**6, 10**6, n).astype(np.int64)
    dk, dv = dev(k), dev(v)
    tbl, rows = statements.groupby_statements([(abi.VH_I32, dk.ptr)], [(abi.VH_I64, dv.ptr)], n)
    keys = tbl.column(0, rows)
    sums = tbl.column(1, rows)
    cnts = tbl.column(2, rows)
    assert rows == len(np.unique(k))
    exp_sum = {}
    for a, b in zip(k.tolist(), v.tolist()):
        exp_sum[a] = exp_sum.get(a, 0) + b
    assert {int(a): s for a, s in zip(keys, sums)} == exp_sum
    assert sum(cnts) == n


@pytest.mark.parametrize("variant", [abi.VH_JOIN01, abi.VH_JOINN])
def test_join_statement_loop(variant):
    print(f"join statement loop variant={variant}", flush=True)
    rng = np.random.default_rng(12)
    nb, n = 30_000, 120_000
    if variant == abi.VH_JOIN01:
        bk = (np.random.default_rng(3).permutation(nb) + 1).astype(np.int32)
    else:
        bk = rng.integers(1, nb // 3, nb).astype(np.int32)   # duplicates
    bp = rng.integers(0, 100, nb).astype(np.int32)
    pk = rng.integers(0, nb + 10, n).astype(np.int32)
    dbk, dbp, dpk = dev(bk), dev(bp), dev(pk)
    pos, pays = statements.join_statements((abi.VH_I32, dbk.ptr), [(abi.VH_I32, dbp.ptr)], nb,
                                           (abi.VH_I32, dpk.ptr), n, variant)
    got = sorted(zip(pos.tolist(), pays[0].tolist()))
    by_key = {}
    for kk, pp in zip(bk.tolist(), bp.tolist()):
        by_key.setdefault(kk, []).append(pp)
    exp = sorted((i, pp) for i, kk in enumerate(pk.tolist()) for pp in by_key.get(kk, []))
    assert got == exp
