"""Generated probe kernels (vh_generate_source): the HIP text for the main
pipeline shapes compiles with hipRTC for gfx950 on the host, for every
group-by strategy, both join layouts and both column-loading orders.  No
GPU needed: lowering and compilation are host work."""
import ctypes as C
import os

import pytest

from vldb_voila_amd import abi, datagen, runtime, workloads
from vldb_voila_amd.pipeline import Pipeline, col, table_from_arrays

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [("vh_fused.hpp", "vldb_voila_amd/csrc/vh_fused.hpp"),
           ("vh_device.hpp", "vldb_voila_amd/csrc/vh_device.hpp"),
           ("vh_plan.hpp", "vldb_voila_amd/csrc/vh_plan.hpp"),
           ("../../include/voila_hip.h", "include/voila_hip.h")]


def _hiprtc():
    try:
        return C.CDLL("/opt/rocm/lib/libhiprtc.so")
    except OSError:
        pytest.skip("hipRTC not installed")


def source(desc, **cfg):
    L = runtime.lib()
    L.vh_generate_source.argtypes = [C.c_void_p, C.c_void_p, C.c_char_p, C.c_size_t, C.POINTER(C.c_size_t)]
    c = runtime.make_config(**cfg)
    n = C.c_size_t()
    runtime.check(L.vh_generate_source(C.byref(desc), C.byref(c), None, 0, C.byref(n)))
    buf = C.create_string_buffer(n.value + 1)
    runtime.check(L.vh_generate_source(C.byref(desc), C.byref(c), buf, n.value + 1, C.byref(n)))
    return buf.value.decode()


def compile_for_gfx950(src):
    R = _hiprtc()
    names = (C.c_char_p * len(HEADERS))(*[n.encode() for n, _ in HEADERS])
    texts = (C.c_char_p * len(HEADERS))(*[open(os.path.join(ROOT, p)).read().encode() for _, p in HEADERS])
    prog = C.c_void_p()
    assert R.hiprtcCreateProgram(C.byref(prog), src.encode(), b"vh_generated.hip", len(HEADERS), texts, names) == 0
    opts = [b"--offload-arch=gfx950", b"-O3", b"-std=c++17", b"-Xclang", b"-target-feature", b"-Xclang",
            b"-packed-fp32-ops"]
    rc = R.hiprtcCompileProgram(prog, len(opts), (C.c_char_p * len(opts))(*opts))
    n = C.c_size_t()
    R.hiprtcGetProgramLogSize(prog, C.byref(n))
    log = C.create_string_buffer(n.value + 1)
    R.hiprtcGetProgramLog(prog, log)
    size = C.c_size_t()
    R.hiprtcGetCodeSize(prog, C.byref(size))
    R.hiprtcDestroyProgram(C.byref(prog))
    return rc, log.value.decode(), size.value


def pipelines():
    b, p = datagen.gen_join_aggregate(2000, 6000, seed=5)
    t = datagen.gen_filter_groupby(2000, seed=1)
    yield "join_aggregate", workloads.join_aggregate_pipeline(b, p).desc()
    yield "join01", workloads.join_aggregate_pipeline(b, p, variant=abi.VH_JOIN01).desc()
    yield "filter_groupby", workloads.filter_groupby_pipeline(t).desc()
    import numpy as np
    rng = np.random.default_rng(1)
    dim = {"k": np.arange(1, 101, dtype=np.int32), "n": rng.integers(0, 7, 100).astype(np.int16)}
    fact = {"a": rng.integers(1, 101, 500).astype(np.int32), "b": rng.integers(1, 101, 500).astype(np.int32),
            "v": rng.integers(-5, 5, 500).astype(np.int8)}
    pl = Pipeline(table_from_arrays("f", fact))
    pl.join(table_from_arrays("d1", dim), "k", "a", payload=["n"], names=["n1"])
    pl.join(table_from_arrays("d2", dim), "k", "b", payload=["n"], names=["n2"])
    pl.groupby(["n1", "n2"], [("sum", col("v") * (col("n1") + 3)), ("count",)])
    yield "two_joins_products", pl.desc()


@pytest.mark.parametrize("name,desc", list(pipelines()))
def test_generated_source_compiles(name, desc):
    for mode in (abi.VH_GROUP_AUTO, abi.VH_GROUP_LDS, abi.VH_GROUP_GLOBAL, abi.VH_GROUP_HASH, abi.VH_GROUP_LDSHASH,
                 abi.VH_GROUP_PART):
        for layout in (abi.VH_LAYOUT_AUTO, abi.VH_LAYOUT_HASH):
            for kernel in (abi.VH_KERNEL_JIT, abi.VH_KERNEL_JIT_LAZY, abi.VH_KERNEL_JIT_PIPE):
                src = source(desc, group_mode=mode, join_layout=layout, kernel=kernel)
                assert 'extern "C" __global__' in src
                rc, log, size = compile_for_gfx950(src)
                assert rc == 0 and size > 0, (name, mode, layout, kernel, log[:2000])


def test_partitioned_hash_table_source():
    """A hash table above 4 MiB is built partitioned: the generated probe
    wraps within the key's sub-table and consults the overflow table after a
    spill; forcing part_shift -1 gives the single-table probe."""
    b, p = datagen.gen_join_sparse(600_000, 6000, seed=5)
    d = workloads.join_aggregate_pipeline(b, p).desc()
    parted = source(d)
    assert "hnext(h[u]" in parted and ".ovf" in parted and ".spill" in parted
    flat = source(d, part_shift=-1)
    assert "hnext(" not in flat and ".ovf" not in flat
    for src in (parted, flat, source(d, part_shift=11, kernel=abi.VH_KERNEL_JIT_PIPE)):
        rc, log, size = compile_for_gfx950(src)
        assert rc == 0 and size > 0, log[:2000]


def split_radix(src):
    i = src.index("// generated join kernel")
    return src[:i], src[i:]


def test_radix_probe_sources():
    """A last join over a partitioned hash table with radix = 1: a probe
    kernel that stages each chunk's tuples in LDS and stores them sorted by
    partition, and a join kernel over the partitions' LDS copies, for every
    group-by strategy it takes, with the carried values (a probe column, an
    earlier join's payload)."""
    import numpy as np
    b, p = datagen.gen_join_sparse(600_000, 6000, seed=5)
    d = workloads.join_aggregate_pipeline(b, p).desc()
    assert "vh_generated_join" not in source(d)
    probe, join = split_radix(source(d, radix=1))
    assert "radix_flush_begin" in probe and "sv0[pos]" in probe and "sv1[pos]" not in probe
    assert "vh_generated_join" in join and "tab[h[u]]" in join and "rv0[x]" in join
    rng = np.random.default_rng(2)
    dim = {"k": np.arange(1, 101, dtype=np.int32), "n": rng.integers(0, 7, 100).astype(np.int16)}
    fact = {"a": rng.integers(1, 101, 500).astype(np.int32), "fk": p["fk"][:500],
            "v": rng.integers(-5, 5, 500).astype(np.int8)}
    pl = Pipeline(table_from_arrays("f", fact))
    pl.join(table_from_arrays("d1", dim), "k", "a", payload=["n"], names=["n1"])
    pl.join(table_from_arrays("r", b), "key", "fk", payload=["segment"], names=["seg"])
    pl.groupby(["n1", "seg"], [("sum", col("v") * (col("n1") + 3)), ("count",)])
    d2 = pl.desc()
    for mode in (abi.VH_GROUP_AUTO, abi.VH_GROUP_GLOBAL, abi.VH_GROUP_HASH, abi.VH_GROUP_LDSHASH):
        for src in (source(d, group_mode=mode, radix=1), source(d2, group_mode=mode, radix=1)):
            probe, join = split_radix(src)
            for part in (probe, join):
                rc, log, size = compile_for_gfx950(part)
                assert rc == 0 and size > 0, (mode, log[:2000])
    probe, join = split_radix(source(d2, radix=1))
    assert "sv1[pos]" in probe and "pay0[u] = rv0[x]" in join


def test_source_reflects_the_plan():
    b, p = datagen.gen_join_aggregate(2000, 6000, seed=5)
    d = workloads.join_aggregate_pipeline(b, p).desc()
    direct = source(d)
    hashed = source(d, join_layout=abi.VH_LAYOUT_HASH)
    # a 2000-key direct table fits in LDS: copied there, looked up there
    assert "P.joins[0].dir" in direct and "td0[off[u]]" in direct
    assert "P.joins[0].slots" in hashed and ".dir" not in hashed and "td0" not in hashed
    # the generic kernel reads no LDS copy, so the source of an unknown
    # kernel choice still names the global table
    big_b, big_p = datagen.gen_join_aggregate(200_000, 6000, seed=5)
    big = source(workloads.join_aggregate_pipeline(big_b, big_p).desc())
    assert "uint8_t*)P.joins[0].dir" in big and "td0" not in big and "tb0" not in big
    # eager: the vector loop loads every column of a chunk before the
    # predicate; the remaining rows take the scalar loop
    eager = source(d)
    assert "q2[p] = __builtin_nontemporal_load" in eager
    assert eager.index("compare(") > eager.index("q2[p] = __builtin_nontemporal_load")
    assert "c2[u] = " in eager
    # lazy: the vector loop loads each chunk's predicate column at the loop
    # top (no pipelining) and the other columns after the predicate, for
    # live rows only (buffer loads); the scalar tail loads them after the
    # predicate too
    lazy = source(d, kernel=abi.VH_KERNEL_JIT_LAZY)
    assert "const i64 pb = (rb >> 1) + w * (32 * U) + lane;" in lazy
    assert "__builtin_amdgcn_raw_buffer_load_b128" in lazy   # int64 price, two rows a lane
    assert lazy.index("compare(") < lazy.index("__builtin_amdgcn_raw_buffer_load")
    tail = lazy[lazy.index("for (i64 base = rb + nfull"):]
    assert tail.index("compare(") < tail.index("c2[u] =")


def test_filtered_dimension_keeps_a_key_bitmap_in_lds():
    """A filtered build too large for an LDS copy of its table keeps a bitmap
    of the keys it holds in LDS; unfiltered dense builds keep none."""
    import numpy as np
    rng = np.random.default_rng(3)
    n = 300_000
    dim = {"k": np.arange(1, n + 1, dtype=np.int32), "r": rng.integers(0, 5, n).astype(np.int32),
           "v": rng.integers(0, 25, n).astype(np.int32)}
    fact = {"fk": rng.integers(1, n + 1, 10_000).astype(np.int32), "m": rng.integers(0, 100, 10_000).astype(np.int64)}
    pl = Pipeline(table_from_arrays("f", fact))
    pl.join(table_from_arrays("d", dim), "k", "fk", payload=["v"], build_preds=[("r", "==", 2)], names=["v"])
    pl.groupby(["v"], [("sum", col("m"))])
    src = source(pl.desc())
    assert "tb0[off[u] >> 5]" in src and "P.joins[0].bits" in src
    assert "td0" not in src


def batch_source(desc, **cfg):
    L = runtime.lib()
    L.vh_generate_batch_source.argtypes = [C.c_void_p, C.c_void_p, C.c_char_p, C.c_size_t, C.POINTER(C.c_size_t)]
    c = runtime.make_config(**cfg)
    n = C.c_size_t()
    runtime.check(L.vh_generate_batch_source(C.byref(desc), C.byref(c), None, 0, C.byref(n)))
    buf = C.create_string_buffer(n.value + 1)
    runtime.check(L.vh_generate_batch_source(C.byref(desc), C.byref(c), buf, n.value + 1, C.byref(n)))
    return buf.value.decode()


@pytest.mark.parametrize("mode", [abi.VH_GROUP_AUTO, abi.VH_GROUP_REG, abi.VH_GROUP_LDS, abi.VH_GROUP_GLOBAL,
                                  abi.VH_GROUP_HASH, abi.VH_GROUP_LDSHASH, abi.VH_GROUP_PART])
def test_batch_entry_point_compiles(mode):
    """The batched launch's generated kernel: work items from a queue, the
    item's rows through the same vector loop, its LDS state flushed."""
    prm = datagen.batch_params(1, seed=3)[0]
    prm["groups"] = 100
    b, p = datagen.gen_batch_pipeline(prm, seed=3, index=0)
    src = batch_source(workloads.batch_pipeline(prm, b, p).desc(), group_mode=mode)
    assert "void vh_generated_batch(" in src and "gatomic_add(counter, 1u)" in src
    assert "const i64 rb = I.r0, re = I.r1;" in src or "const i64 first = I.r0;" in src
    if mode in (abi.VH_GROUP_LDS, abi.VH_GROUP_REG):
        assert "lds_dense_flush(P, L);" in src
    rc, log, size = compile_for_gfx950(src)
    assert rc == 0 and size > 0, log[:2000]


def test_mixed_sign_product_is_refused():
    """The reference types a product by its operands' (min*min, max*max)
    (typing_pass.cpp:164-176).  With one operand of mixed sign that interval
    misses the true one: a in [-60, 1] times b in [1, 3] is typed i8 over
    [-60, 3] while the values span [-180, 3], so the reference's vector
    flavor would wrap.  The lowering refuses such a sum instead of producing
    wrapped numbers; the same product with both operands non-negative is
    accepted."""
    import numpy as np
    rng = np.random.default_rng(3)
    n = 1000
    a = rng.integers(-60, 2, n).astype(np.int32)
    b = rng.integers(1, 4, n).astype(np.int32)
    a[:2], b[:2] = (-60, 1), (1, 3)
    g = rng.integers(0, 3, n).astype(np.int32)
    t = table_from_arrays("t", {"a": a, "b": b, "g": g})
    bad = Pipeline(t).groupby(["g"], [("sum", col("a") * col("b"))]).desc()
    L = runtime.lib()
    L.vh_generate_source.argtypes = [C.c_void_p, C.c_void_p, C.c_char_p, C.c_size_t, C.POINTER(C.c_size_t)]
    n_out = C.c_size_t()
    cfg = runtime.make_config()
    rc = L.vh_generate_source(C.byref(bad), C.byref(cfg), None, 0, C.byref(n_out))
    assert rc == 2   # VH_ERR_UNSUPPORTED
    assert b"may wrap its type" in L.vh_last_error()
    t2 = table_from_arrays("t", {"a": a + 60, "b": b, "g": g})
    good = Pipeline(t2).groupby(["g"], [("sum", col("a") * col("b"))]).desc()
    assert "vh_generated" in source(good)
