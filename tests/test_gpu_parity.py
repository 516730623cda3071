"""GPU parity: the fused gfx950 path (through the C ABI) against the CPU
oracle, bit-exact on sorted QueryResult rows."""
import numpy as np
import pytest

import oracle
from vldb_voila_amd import abi, datagen, runtime, workloads
from vldb_voila_amd.pipeline import Pipeline, col, table_from_arrays

pytestmark = pytest.mark.gpu

MODES = [abi.VH_GROUP_AUTO, abi.VH_GROUP_LDS, abi.VH_GROUP_GLOBAL, abi.VH_GROUP_HASH,
         abi.VH_GROUP_LDSHASH, abi.VH_GROUP_PART]


def gpu_rows(builder_fn, mode=abi.VH_GROUP_AUTO, **cfg):
    """builder_fn(addr_map) -> Pipeline over device copies."""
    return oracle.sorted_rows(runtime.run_pipeline(builder_fn(), runtime.make_config(group_mode=mode, **cfg)))


def on_gpu(arrays_list, make):
    devs = [runtime.upload(a) for a in arrays_list]
    return make(*[runtime.addrs(d) for d in devs]), devs


@pytest.fixture(scope="module", autouse=True)
def device():
    print("gpu parity: init device 0", flush=True)
    runtime.init(0)


@pytest.mark.parametrize("kernel", [abi.VH_KERNEL_GENERIC, abi.VH_KERNEL_JIT, abi.VH_KERNEL_JIT_PIPE])
@pytest.mark.parametrize("mode", [abi.VH_GROUP_AUTO, abi.VH_GROUP_REG] + MODES[1:])
def test_filter_groupby_small(mode, kernel):
    print(f"filter_groupby 300k mode={mode} kernel={kernel}", flush=True)
    t = datagen.gen_filter_groupby(300_000, seed=11)
    want = oracle.sorted_rows(oracle.bruteforce(workloads.filter_groupby_pipeline(t).desc()))
    p, devs = on_gpu([t], lambda a: workloads.filter_groupby_pipeline(t, a))
    got = oracle.sorted_rows(runtime.run_pipeline(p, runtime.make_config(group_mode=mode, kernel=kernel)))
    assert got == want


KERNELS = [abi.VH_KERNEL_GENERIC, abi.VH_KERNEL_JIT, abi.VH_KERNEL_JIT_LAZY, abi.VH_KERNEL_JIT_PIPE]


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("mode", MODES)
def test_join_aggregate_small(mode, kernel):
    print(f"join_aggregate 50k/200k mode={mode} kernel={kernel}", flush=True)
    b, pr = datagen.gen_join_aggregate(50_000, 200_000, seed=13)
    want = oracle.sorted_rows(oracle.bruteforce(workloads.join_aggregate_pipeline(b, pr).desc()))
    p, devs = on_gpu([b, pr], lambda x, y: workloads.join_aggregate_pipeline(b, pr, x, y))
    got = oracle.sorted_rows(runtime.run_pipeline(p, runtime.make_config(group_mode=mode, kernel=kernel)))
    assert got == want


def test_generated_kernel_is_the_default():
    """vh_query_new compiles the generated kernel; its source names the
    plan's choices (direct byte table for the dense join)."""
    print("generated kernel in use", flush=True)
    b, pr = datagen.gen_join_aggregate(5_000, 20_000, seed=2)
    p, devs = on_gpu([b, pr], lambda x, y: workloads.join_aggregate_pipeline(b, pr, x, y))
    q = runtime.Query(p.desc())
    kernel, src = q.kernel()
    assert kernel == abi.VH_KERNEL_JIT
    assert "vh_generated" in src and "uint8_t*)P.joins[0].dir" in src
    q.run()
    assert oracle.sorted_rows(q.result()) == oracle.sorted_rows(
        oracle.bruteforce(workloads.join_aggregate_pipeline(b, pr).desc()))
    g = runtime.Query(p.desc(), runtime.make_config(kernel=abi.VH_KERNEL_GENERIC))
    assert g.kernel() == (abi.VH_KERNEL_GENERIC, "")


def test_filter_groupby_full_size():
    print("filter_groupby 6M rows seed 42", flush=True)
    t = datagen.gen_filter_groupby(6_000_000, seed=42)
    d = workloads.filter_groupby_pipeline(t).desc()
    want = oracle.sorted_rows(oracle.run_pipeline(d, threads=8)[0])
    p, devs = on_gpu([t], lambda a: workloads.filter_groupby_pipeline(t, a))
    q = runtime.Query(p.desc())
    for rep in range(3):     # repeated runs reuse the tables and the merge counter
        q.run()
        assert oracle.sorted_rows(q.result()) == want
    # size-independent properties: counts add up to the selected rows
    m = t["date"] <= workloads.FILTER_GROUPBY_CUTOFF
    rows = [r.split("|") for r in want]
    assert sum(int(r[5]) for r in rows) == int(m.sum())
    assert len(rows) == 6


def test_join_aggregate_full_size():
    print("join_aggregate 1.5M/6M seed 42", flush=True)
    b, pr = datagen.gen_join_aggregate(1_500_000, 6_000_000, seed=42)
    want = oracle.sorted_rows(oracle.bruteforce(workloads.join_aggregate_pipeline(b, pr).desc()))
    p, devs = on_gpu([b, pr], lambda x, y: workloads.join_aggregate_pipeline(b, pr, x, y))
    q = runtime.Query(p.desc())
    for rep in range(3):
        q.run()
        assert oracle.sorted_rows(q.result()) == want
    m = pr["date"] <= workloads.JOIN_DATE_CUTOFF
    assert sum(int(r.split("|")[1]) for r in want) == int(pr["price"][m].sum())


def test_joinn_duplicate_keys_last_join():
    print("JoinN duplicates", flush=True)
    rng = np.random.default_rng(5)
    build = {"key": rng.integers(0, 2000, 6000).astype(np.int32),
             "pay": rng.integers(-7, 9, 6000).astype(np.int32)}
    probe = {"fk": rng.integers(-10, 2100, 50000).astype(np.int32),
             "v": rng.integers(-10**9, 10**9, 50000).astype(np.int64)}

    def make(bb, pp):
        p = Pipeline(table_from_arrays("s", probe, pp))
        p.join(table_from_arrays("r", build, bb), "key", "fk", payload=["pay"],
               variant=abi.VH_JOINN, names=["pay"])
        p.groupby(["pay"], [("sum", col("v") * 3), ("count",)])
        return p
    want = oracle.sorted_rows(oracle.bruteforce(make(None, None).desc()))
    p, devs = on_gpu([build, probe], make)
    for mode in MODES:
        got = oracle.sorted_rows(runtime.run_pipeline(p, runtime.make_config(group_mode=mode)))
        assert got == want, mode


def test_join01_first_match_unique_keys():
    print("Join01 unique keys", flush=True)
    b, pr = datagen.gen_join_aggregate(10_000, 80_000, seed=3)
    want = oracle.sorted_rows(oracle.bruteforce(
        workloads.join_aggregate_pipeline(b, pr, variant=abi.VH_JOIN01).desc()))
    p, devs = on_gpu([b, pr], lambda x, y: workloads.join_aggregate_pipeline(b, pr, x, y, variant=abi.VH_JOIN01))
    assert oracle.sorted_rows(runtime.run_pipeline(p)) == want


def test_wide_sums_signed_i64_into_i128():
    print("wide signed sums", flush=True)
    rng = np.random.default_rng(9)
    n = 400_000
    t = {"g": rng.integers(0, 5, n).astype(np.int32),
         "a": rng.integers(-2**62, 2**62, n).astype(np.int64),
         "b": rng.integers(-2**40, 2**40, n).astype(np.int64)}

    def make(addr):
        p = Pipeline(table_from_arrays("t", t, addr))
        p.groupby(["g"], [("sum", col("a")), ("sum", col("a") - col("b")), ("count",)])
        return p
    want = oracle.sorted_rows(oracle.bruteforce(make(None).desc()))
    p, devs = on_gpu([t], make)
    for mode in [abi.VH_GROUP_AUTO] + MODES:
        got = oracle.sorted_rows(runtime.run_pipeline(p, runtime.make_config(group_mode=mode)))
        assert got == want, mode


def test_many_groups_hash_and_global():
    print("many groups", flush=True)
    rng = np.random.default_rng(21)
    n = 500_000
    t = {"k1": rng.integers(0, 700, n).astype(np.int32),
         "k2": rng.integers(-50, 50, n).astype(np.int32),
         "v": rng.integers(0, 10**6, n).astype(np.int64)}

    def make(addr):
        p = Pipeline(table_from_arrays("t", t, addr))
        p.select("v", ">=", 1000)
        p.groupby(["k1", "k2"], [("sum", col("v")), ("count",)])
        return p
    want = oracle.sorted_rows(oracle.bruteforce(make(None).desc()))
    p, devs = on_gpu([t], make)
    for mode in (abi.VH_GROUP_AUTO, abi.VH_GROUP_GLOBAL, abi.VH_GROUP_HASH, abi.VH_GROUP_LDSHASH,
                 abi.VH_GROUP_PART):
        for kernel in (abi.VH_KERNEL_GENERIC, abi.VH_KERNEL_JIT, abi.VH_KERNEL_JIT_PIPE):
            got = oracle.sorted_rows(runtime.run_pipeline(p, runtime.make_config(group_mode=mode, kernel=kernel)))
            assert got == want, (mode, kernel)


@pytest.mark.parametrize("rows", [0, 1, 1023, 300_001])
def test_partitioned_group_by_shapes(rows):
    """GB_PART over a 1M-group domain: regions (workgroup slabs) of uneven
    size, partitions with no tuples, JoinN fan-out past a region's tuple
    space (added with atomics), signed i64 sums into i128, grids that leave
    the last region short."""
    print(f"partitioned group-by rows={rows}", flush=True)
    rng = np.random.default_rng(rows + 5)
    build = {"key": rng.integers(0, 3000, 9000).astype(np.int32),        # ~3 rows per key
             "pay": rng.integers(0, 7, 9000).astype(np.int32)}
    probe = {"fk": rng.integers(0, 3100, rows).astype(np.int32),
             "g": rng.integers(0, 1_000_000, rows).astype(np.int32),
             "v": rng.integers(-2**62, 2**62, rows).astype(np.int64),
             "w": rng.integers(0, 2**20, rows).astype(np.int32)}

    def make(bb, pp):
        p = Pipeline(table_from_arrays("s", probe, pp))
        p.join(table_from_arrays("r", build, bb), "key", "fk", payload=["pay"], variant=abi.VH_JOINN, names=["pay"])
        p.groupby(["g"], [("sum", col("v")), ("sum", col("w") * col("pay")), ("count",)])
        return p
    want = oracle.sorted_rows(oracle.bruteforce(make(None, None).desc()))
    p, devs = on_gpu([build, probe], make)
    for kernel in (abi.VH_KERNEL_GENERIC, abi.VH_KERNEL_JIT, abi.VH_KERNEL_JIT_PIPE):
        for grid in (0, 3, 64):
            cfg = runtime.make_config(group_mode=abi.VH_GROUP_PART, kernel=kernel, grid=grid)
            assert oracle.sorted_rows(runtime.run_pipeline(p, cfg)) == want, (kernel, grid)


def test_global_aggregate_and_empty_inputs():
    print("global aggregate / empty", flush=True)
    t = datagen.gen_filter_groupby(100_000, seed=2)

    def make(addr, cutoff):
        p = Pipeline(table_from_arrays("t", t, addr))
        p.select("date", "<=", cutoff)
        p.groupby([], [("sum", col("price") * col("qty")), ("count",)])
        return p
    for cutoff in (workloads.FILTER_GROUPBY_CUTOFF, 0):
        want = oracle.sorted_rows(oracle.bruteforce(make(None, cutoff).desc()))
        p, devs = on_gpu([t], lambda a: make(a, cutoff))
        assert oracle.sorted_rows(runtime.run_pipeline(p)) == want
    # empty probe and empty build
    e = {"k": np.zeros(0, dtype=np.int32), "v": np.zeros(0, dtype=np.int64)}
    p, devs = on_gpu([e], lambda a: Pipeline(table_from_arrays("e", e, a)).groupby(["k"], [("sum", col("v"))]))
    assert runtime.run_pipeline(p) == ""
    b, pr = datagen.gen_join_aggregate(1000, 5000, seed=1)
    b0 = {k: v[:0] for k, v in b.items()}
    p, devs = on_gpu([b0, pr], lambda x, y: workloads.join_aggregate_pipeline(b0, pr, x, y))
    assert runtime.run_pipeline(p) == ""


def test_build_side_filter_two_joins():
    print("two joins with build filters", flush=True)
    rng = np.random.default_rng(17)
    cust = {"ck": np.arange(1, 30001, dtype=np.int32), "region": rng.integers(0, 5, 30000).astype(np.int32),
            "nation": rng.integers(0, 25, 30000).astype(np.int32)}
    supp = {"sk": np.arange(1, 2001, dtype=np.int32), "region": rng.integers(0, 5, 2000).astype(np.int32),
            "nation": rng.integers(0, 25, 2000).astype(np.int32)}
    fact = {"c": rng.integers(1, 30001, 300_000).astype(np.int32),
            "s": rng.integers(1, 2001, 300_000).astype(np.int32),
            "rev": rng.integers(100000, 10**7, 300_000).astype(np.int64),
            "cost": rng.integers(0, 10**6, 300_000).astype(np.int64)}

    def make(fa, ca, sa):
        p = Pipeline(table_from_arrays("f", fact, fa))
        p.join(table_from_arrays("c", cust, ca), "ck", "c", payload=["nation"], build_preds=[("region", "==", 2)],
               names=["c_nation"])
        p.join(table_from_arrays("s", supp, sa), "sk", "s", payload=["nation"], build_preds=[("region", "==", 2)],
               names=["s_nation"])
        p.groupby(["c_nation", "s_nation"], [("sum", col("rev") - col("cost")), ("count",)])
        return p
    want = oracle.sorted_rows(oracle.bruteforce(make(None, None, None).desc()))
    p, devs = on_gpu([fact, cust, supp], make)
    for mode in MODES:
        got = oracle.sorted_rows(runtime.run_pipeline(p, runtime.make_config(group_mode=mode)))
        assert got == want, mode


def test_lds_key_bitmap_with_repeated_keys():
    """A filtered build too large for an LDS copy of its table: the generated
    kernel tests a key bitmap in LDS before the table in global memory; keys
    repeated in the build (JoinN, last join) still reach every row."""
    print("LDS key bitmap, filtered build with repeated keys", flush=True)
    rng = np.random.default_rng(23)
    n = 400_000
    keys = np.arange(1, n + 1, dtype=np.int32)
    keys[rng.integers(0, n, 5000)] = rng.integers(1, n + 1, 5000).astype(np.int32)   # repeats
    dim = {"k": keys, "r": rng.integers(0, 4, n).astype(np.int32), "v": rng.integers(0, 9, n).astype(np.int32)}
    fact = {"fk": rng.integers(-5, n + 10, 500_000).astype(np.int32),
            "m": rng.integers(-10**6, 10**6, 500_000).astype(np.int64)}

    def make(fa, da):
        p = Pipeline(table_from_arrays("f", fact, fa))
        p.join(table_from_arrays("d", dim, da), "k", "fk", payload=["v"], build_preds=[("r", "==", 1)],
               variant=abi.VH_JOINN, names=["v"])
        p.groupby(["v"], [("sum", col("m")), ("count",)])
        return p
    want = oracle.sorted_rows(oracle.bruteforce(make(None, None).desc()))
    p, devs = on_gpu([fact, dim], make)
    q = runtime.Query(p.desc())
    kernel, src = q.kernel()
    assert kernel == abi.VH_KERNEL_JIT and "tb0[off[u] >> 5]" in src
    for _ in range(2):   # the bitmap is rebuilt with the table on every run
        q.run()
        assert oracle.sorted_rows(q.result()) == want
    q.free()


def test_key_outside_declared_range_is_an_error():
    print("bad metadata", flush=True)
    t = {"k": np.array([1, 2, 3, 9], dtype=np.int32), "v": np.array([1, 2, 3, 4], dtype=np.int64)}
    p, devs = on_gpu([t], lambda a: Pipeline(table_from_arrays("t", t, a)).groupby(["k"], [("sum", col("v"))]))
    p.probe.columns[0].vmax = 3   # lie about max
    with pytest.raises(runtime.VoilaHipError):
        runtime.run_pipeline(p)


LAYOUTS = [abi.VH_LAYOUT_HASH, abi.VH_LAYOUT_DIRECT, abi.VH_LAYOUT_AUTO]


@pytest.mark.parametrize("prange", [200, 60_000, 3_000_000_000])
def test_join_layouts_payload_widths(prange):
    """Direct tables hold the payload in 1, 2 or 4 bytes (by its range);
    repeated keys spill into the overflow slots.  Every layout and group-by
    mode gives the oracle's rows, for JoinN with duplicates and Join01."""
    print(f"join layouts payload range {prange}", flush=True)
    rng = np.random.default_rng(prange % 1000 + 1)
    nb, n = 8000, 60_000
    for variant in (abi.VH_JOINN, abi.VH_JOIN01):
        if variant == abi.VH_JOINN:
            bk = rng.integers(10, 5000, nb).astype(np.int32)           # repeated keys
        else:
            bk = (rng.permutation(nb) * 3 + 7).astype(np.int32)       # unique, sparse
        build = {"k": bk, "p": rng.integers(0, prange, nb).astype(np.int64)}
        probe = {"fk": rng.integers(0, 3 * nb + 20, n).astype(np.int32),
                 "v": rng.integers(-1000, 1000, n).astype(np.int64)}

        def make(bb, pp):
            p = Pipeline(table_from_arrays("s", probe, pp))
            p.join(table_from_arrays("r", build, bb), "k", "fk", payload=["p"], variant=variant, names=["p"])
            p.groupby(["p"] if prange < 100_000 else [], [("sum", col("v")), ("count",)])
            return p
        want = oracle.sorted_rows(oracle.bruteforce(make(None, None).desc()))
        p, devs = on_gpu([build, probe], make)
        for layout in LAYOUTS:
            for mode in (abi.VH_GROUP_AUTO, abi.VH_GROUP_GLOBAL, abi.VH_GROUP_HASH):
                got = oracle.sorted_rows(runtime.run_pipeline(
                    p, runtime.make_config(group_mode=mode, join_layout=layout)))
                assert got == want, (variant, layout, mode)


def test_direct_overflow_cleared_between_runs():
    """One query, build data rewritten in place between runs: repeated keys,
    then unique keys, then repeated keys again.  The overflow slots a run
    filled must not leak matches into the next one."""
    print("direct layout overflow reuse", flush=True)
    rng = np.random.default_rng(21)
    nb, n = 4000, 40_000
    dup_keys = rng.integers(0, 1000, nb).astype(np.int32)
    uniq_keys = rng.permutation(nb).astype(np.int32)
    dup_keys[0], dup_keys[1] = 0, nb - 1            # same [min, max] as the unique keys
    build = {"k": dup_keys.copy(), "p": rng.integers(0, 50, nb).astype(np.int32)}
    probe = {"fk": rng.integers(0, nb, n).astype(np.int32), "v": rng.integers(0, 100, n).astype(np.int64)}
    bdev, pdev = runtime.upload(build), runtime.upload(probe)

    def make(keys, bb=None, pp=None):
        b = dict(build, k=keys)
        p = Pipeline(table_from_arrays("s", probe, pp))
        p.join(table_from_arrays("r", b, bb), "k", "fk", payload=["p"], variant=abi.VH_JOINN, names=["p"])
        p.groupby(["p"], [("sum", col("v")), ("count",)])
        return p
    q = runtime.Query(make(dup_keys, runtime.addrs(bdev), runtime.addrs(pdev)).desc(),
                      runtime.make_config(join_layout=abi.VH_LAYOUT_DIRECT))
    for keys in (dup_keys, uniq_keys, dup_keys, uniq_keys):
        bdev["k"].copy_from(keys)
        q.run()
        assert oracle.sorted_rows(q.result()) == oracle.sorted_rows(oracle.bruteforce(make(keys).desc()))


@pytest.mark.parametrize("merge", [1, 2])
def test_dense_group_by_merge_modes(merge):
    """Workgroups of the dense LDS group-by either leave partials for a merge
    launch (1) or add into the result with 128-bit atomics (2); both are
    exact, for both kernels and the lane-private and shared layouts."""
    print(f"dense group-by merge={merge}", flush=True)
    t = datagen.gen_filter_groupby(200_000, seed=31)
    b, pr = datagen.gen_join_aggregate(20_000, 100_000, seed=32)
    cases = [(workloads.filter_groupby_pipeline(t).desc(), [t],
              lambda a: workloads.filter_groupby_pipeline(t, a)),
             (workloads.join_aggregate_pipeline(b, pr).desc(), [b, pr],
              lambda x, y: workloads.join_aggregate_pipeline(b, pr, x, y))]
    for desc, arrays, make in cases:
        want = oracle.sorted_rows(oracle.bruteforce(desc))
        p, devs = on_gpu(arrays, make)
        for mode in (abi.VH_GROUP_REG, abi.VH_GROUP_LDS):
            for kernel in (abi.VH_KERNEL_GENERIC, abi.VH_KERNEL_JIT):
                for grid in (0, 37):
                    cfg = runtime.make_config(group_mode=mode, kernel=kernel, merge=merge, grid=grid)
                    got = oracle.sorted_rows(runtime.run_pipeline(p, cfg))
                    assert got == want, (mode, kernel, grid)


def test_unroll_variants_of_the_generated_kernel():
    print("generated kernel unroll 1..16", flush=True)
    b, pr = datagen.gen_join_aggregate(20_000, 100_003, seed=33)
    want = oracle.sorted_rows(oracle.bruteforce(workloads.join_aggregate_pipeline(b, pr).desc()))
    p, devs = on_gpu([b, pr], lambda x, y: workloads.join_aggregate_pipeline(b, pr, x, y))
    for unroll in (1, 2, 8, 16):
        for kernel in (abi.VH_KERNEL_JIT, abi.VH_KERNEL_JIT_PIPE):
            cfg = runtime.make_config(kernel=kernel, unroll=unroll)
            assert oracle.sorted_rows(runtime.run_pipeline(p, cfg)) == want, (unroll, kernel)


@pytest.mark.parametrize("theta", [0.0, 1.0])
def test_star_schema(theta):
    """Four joins (two dimensions filtered on region, part a pure semi-join),
    175 groups; uniform and Zipf(1.0) part keys; every generated-kernel
    variant and the generic kernel."""
    print(f"star schema 2M fact rows theta={theta}", flush=True)
    fact, date, cust, supp, part = datagen.gen_star(2_000_000, zipf_part=theta, seed=7, dims_scale=0.2)
    want = oracle.sorted_rows(oracle.run_pipeline(
        workloads.star_pipeline(fact, date, cust, supp, part).desc(), threads=8)[0])
    assert len(want) == 175
    dev = {k: runtime.upload(v) for k, v in
           (("fact", fact), ("date", date), ("customer", cust), ("supplier", supp), ("part", part))}
    addrs = {k: runtime.addrs(v) for k, v in dev.items()}
    p = workloads.star_pipeline(fact, date, cust, supp, part, addrs)
    for kernel in KERNELS:
        got = oracle.sorted_rows(runtime.run_pipeline(p, runtime.make_config(kernel=kernel)))
        assert got == want, kernel


def voila_hash_np(x: np.ndarray) -> np.ndarray:
    """voila_hash over uint64 (runtime.hpp:413-424), wrapping arithmetic."""
    x = x.astype(np.uint64)
    m = np.uint64(0x45d9f3b)
    x = x ^ (x >> np.uint64(16))
    x = x * m
    x = x ^ (x >> np.uint64(16))
    x = x * m
    return x ^ (x >> np.uint64(16))


def colliding_offsets(count, bits=16):
    """Key offsets o (0 first) whose voila_hash(o) has `bits` low zero bits:
    they share one slot of any table of up to 2^bits slots."""
    out = [0]
    start = 1
    while len(out) < count:
        o = np.arange(start, start + (1 << 22), dtype=np.uint64)
        out.extend(o[(voila_hash_np(o) & np.uint64((1 << bits) - 1)) == 0].tolist())
        start += 1 << 22
    return np.array(out[:count], dtype=np.int64)


def test_colliding_hash_keys_near_int64_max():
    """400 build keys near 2^62 whose table offsets all hash to one slot:
    every probe walks one long cluster (hits, and colliding misses that walk
    it to the end); sums of values near +-2^61 need 128 bits.  A key range
    wider than 32 bits is refused."""
    print("colliding hash keys", flush=True)
    rng = np.random.default_rng(77)
    offs = colliding_offsets(700)
    base = np.int64(2**62)
    build = {"k": base + offs[:400], "p": rng.integers(0, 40, 400).astype(np.int64)}
    n = 50_000
    pick = rng.integers(0, 3, n)
    fk = np.where(pick == 0, base + offs[rng.integers(0, 400, n)],
                  np.where(pick == 1, base + offs[400 + rng.integers(0, 300, n)],
                           base + rng.integers(0, int(offs.max()), n)))
    probe = {"fk": fk.astype(np.int64),
             "v": rng.integers(-(2**61), 2**61, n).astype(np.int64)}
    for variant in (abi.VH_JOINN, abi.VH_JOIN01):
        def make(bb, pp):
            p = Pipeline(table_from_arrays("s", probe, pp))
            p.join(table_from_arrays("r", build, bb), "k", "fk", payload=["p"], variant=variant, names=["p"])
            p.groupby(["p"], [("sum", col("v")), ("count",)])
            return p
        want = oracle.sorted_rows(oracle.bruteforce(make(None, None).desc()))
        assert len(want) == 40
        p, devs = on_gpu([build, probe], make)
        for kernel in KERNELS:
            for layout in (abi.VH_LAYOUT_AUTO, abi.VH_LAYOUT_HASH):
                got = oracle.sorted_rows(runtime.run_pipeline(
                    p, runtime.make_config(kernel=kernel, join_layout=layout)))
                assert got == want, (variant, kernel, layout)
    wide = {"k": np.array([0, 2**40], dtype=np.int64), "p": np.array([1, 2], dtype=np.int64)}
    p = Pipeline(table_from_arrays("s", probe))
    p.join(table_from_arrays("r", wide), "k", "fk", payload=["p"], names=["p"])
    p.groupby(["p"], [("count",)])
    with pytest.raises(runtime.VoilaHipError):
        runtime.Query(p.desc())


def test_colliding_group_keys():
    """600 group keys whose voila_hash shares its low 16 bits (so they share
    one slot of the global hash table's first probe, and of the LDS hash
    table's), 30 K rows: every hash group-by strategy gives the oracle's rows."""
    print("colliding group keys", flush=True)
    rng = np.random.default_rng(78)
    keys = colliding_offsets(600)
    t = {"g": keys[rng.integers(0, 600, 30_000)], "v": rng.integers(-10**12, 10**12, 30_000)}

    def make(addrs):
        p = Pipeline(table_from_arrays("t", t, addrs))
        p.groupby(["g"], [("sum", col("v")), ("count",)])
        return p
    want = oracle.sorted_rows(oracle.bruteforce(make(None).desc()))
    assert len(want) == 600
    p, devs = on_gpu([t], make)
    for mode in (abi.VH_GROUP_AUTO, abi.VH_GROUP_HASH, abi.VH_GROUP_LDSHASH):
        for kernel in (abi.VH_KERNEL_GENERIC, abi.VH_KERNEL_JIT):
            got = oracle.sorted_rows(runtime.run_pipeline(p, runtime.make_config(group_mode=mode, kernel=kernel)))
            assert got == want, (mode, kernel)


def test_more_than_2_31_rows():
    """2.3 G rows of a u8 column (row indices past 2^31): select x <= 100,
    sum(x) and count(*), against the closed form of the repeating pattern
    x = i mod 251."""
    print("2.3G rows", flush=True)
    n, period = 2_300_000_000, 251
    pattern = (np.arange(period) % 256).astype(np.uint8)
    x = np.tile(pattern, n // period + 1)[:n]
    full, rest = divmod(n, period)
    want_cnt = full * 101 + sum(1 for v in range(rest) if v <= 100)
    want_sum = full * sum(range(101)) + sum(v for v in range(rest) if v <= 100)
    dev = runtime.DeviceBuffer.from_numpy(x)
    t = table_from_arrays("t", {"x": x}, {"x": dev.ptr}, {"x": (0, 250)})
    del x
    p = Pipeline(t)
    p.select("x", "<=", 100)
    p.groupby([], [("sum", col("x")), ("count",)])
    for kernel in (abi.VH_KERNEL_JIT, abi.VH_KERNEL_GENERIC):
        got = runtime.run_pipeline(p, runtime.make_config(kernel=kernel)).strip()
        assert got == f"{want_sum}|{want_cnt}", (kernel, got)


def test_global_aggregate_register_sums():
    """One group (no keys): every lane sums in registers and adds into LDS
    once.  Signed 64-bit values near 2^62 into an int128 sum (the carry
    count), a narrow signed sum and a count, with and without a selection,
    as a single query and as a batch's pipeline."""
    print("global aggregate, register sums", flush=True)
    rng = np.random.default_rng(17)
    n = 3_000_000
    t = {"a": rng.integers(-2**62, 2**62, n).astype(np.int64),
         "c": rng.integers(-30000, 30000, n).astype(np.int32),
         "f": rng.integers(0, 10, n).astype(np.int32)}

    def make(addr, select):
        p = Pipeline(table_from_arrays("t", t, addr))
        if select:
            p.select("f", "<", 7)
        p.groupby([], [("sum", col("a")), ("sum", col("c")), ("count",)])
        return p
    devs = [runtime.upload(t)]
    for select in (False, True):
        want = oracle.sorted_rows(oracle.bruteforce(make(None, select).desc()))
        p = make(runtime.addrs(devs[0]), select)
        src = runtime.Query(p.desc()).kernel()[1]
        assert "racc_lo" in src
        got = oracle.sorted_rows(runtime.run_pipeline(p))
        assert got == want, select
        batch = runtime.Batch([p.desc(), p.desc()])
        s = runtime.Stream()
        for rep in range(2):
            batch.run_graph(s)
            s.synchronize()
            for i in range(2):
                assert oracle.sorted_rows(batch.result(i, s)) == want, (select, rep, i)
        batch.free()


def test_partitioned_group_by_packed_records():
    """Partitioned group-by records: a signed argument that fits 32 bits is
    kept in 4 bytes and sign-extended, a wide one in 8 (into int128 sums),
    beside the group index; single query and batch, repeated runs (the
    result arrays are not cleared between runs without a fan-out)."""
    print("partitioned group-by, packed records", flush=True)
    rng = np.random.default_rng(23)
    n = 1_500_000
    t = {"g": rng.integers(0, 200_000, n).astype(np.int32),
         "c": rng.integers(-2**31, 2**31 - 1, n).astype(np.int32),
         "a": rng.integers(-2**62, 2**62, n).astype(np.int64),
         "f": rng.integers(0, 4, n).astype(np.int32)}

    def make(addr):
        p = Pipeline(table_from_arrays("t", t, addr))
        p.select("f", "!=", 2)
        p.groupby(["g"], [("sum", col("c")), ("sum", col("a")), ("count",)])
        return p
    want = oracle.sorted_rows(oracle.bruteforce(make(None).desc()))
    devs = [runtime.upload(t)]
    p = make(runtime.addrs(devs[0]))
    cfg = runtime.make_config(group_mode=abi.VH_GROUP_PART)
    q = runtime.Query(p.desc(), cfg)
    assert q.info()["group_mode"] == abi.VH_GROUP_PART
    for rep in range(2):
        q.run()
        assert oracle.sorted_rows(q.result()) == want, rep
    batch = runtime.Batch([p.desc()], [cfg])
    s = runtime.Stream()
    for rep in range(3):
        batch.run_graph(s)
        s.synchronize()
        assert oracle.sorted_rows(batch.result(0, s)) == want, rep
    batch.free()
