"""VOILA statement sequences executed one primitive launch at a time.

This is the statement-level path: the same loops relalg_translator.cpp emits
for HashAggr (:582-656, the miss/hit chain walk with bucket_insert retries)
and HashJoin (:828-1111, materialise / bucket_build / probe), each statement
a vh_vec_* kernel over whole device columns.  The fused kernels compute the
same results in one pass; this path exists for programs the fusion does not
cover and as a statement-by-statement cross-check.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import abi
from .runtime import DeviceBuffer, check, init, lib


def _bind():
    L = lib()
    vp, i32, i64 = C.c_void_p, C.c_int32, C.c_int64
    L.vh_table_new.argtypes = [C.POINTER(vp), i32, vp, i64]
    L.vh_table_free.argtypes = [vp]
    L.vh_table_reset.argtypes = [vp, vp]
    L.vh_table_rows.argtypes = [vp, vp, C.POINTER(i64)]
    L.vh_table_column.argtypes = [vp, i32, C.POINTER(vp)]
    L.vh_vec_seltrue.argtypes = [vp, i64, vp, vp, vp, vp]
    L.vh_vec_selfalse.argtypes = [vp, i64, vp, vp, vp, vp]
    L.vh_vec_cmp.argtypes = [i32, i32, vp, i64, vp, vp, vp, i64, vp]
    L.vh_vec_arith.argtypes = [i32, i32, vp, i64, vp, i32, vp, i32, vp, i64, vp]
    L.vh_vec_hash.argtypes = [i32, vp, i64, vp, vp, vp]
    L.vh_vec_rehash.argtypes = [i32, vp, i64, vp, vp, vp, vp]
    for n in ("vh_vec_bucket_lookup", "vh_vec_bucket_next", "vh_vec_bucket_insert"):
        getattr(L, n).argtypes = [vp, vp, i64, vp, vp, vp]
    L.vh_vec_check.argtypes = [vp, i32, vp, i64, vp, vp, i32, vp, vp]
    L.vh_vec_gather.argtypes = [vp, i32, vp, i64, vp, i32, vp, vp]
    L.vh_vec_scatter.argtypes = [vp, i32, vp, i64, vp, i32, vp, vp]
    L.vh_vec_aggr_sum.argtypes = [vp, i32, vp, i64, vp, i32, vp, vp]
    L.vh_vec_aggr_count.argtypes = [vp, i32, vp, i64, vp, vp]
    L.vh_vec_write_pos.argtypes = [vp, i64, C.POINTER(i64)]
    L.vh_vec_write.argtypes = [vp, i32, i64, vp, i64, i32, vp, vp]
    L.vh_vec_bucket_build.argtypes = [vp, vp]
    return L


class Table:
    """Device chained hash table (vh_table)."""

    def __init__(self, ncols, wide, max_rows):
        init()
        self.L = _bind()
        w = (C.c_int32 * max(ncols, 1))(*wide)
        p = C.c_void_p()
        check(self.L.vh_table_new(C.byref(p), ncols, w, max_rows))
        self.ptr = p.value
        self.ncols = ncols
        self.wide = list(wide)

    def rows(self):
        n = C.c_int64()
        check(self.L.vh_table_rows(self.ptr, None, C.byref(n)))
        return n.value

    def column(self, c, n):
        p = C.c_void_p()
        check(self.L.vh_table_column(self.ptr, c, C.byref(p)))
        words = 2 if self.wide[c] else 1
        out = np.empty(n * words, dtype=np.uint64)
        if n:
            check(lib().vh_memcpy_d2h(out.ctypes.data, p.value, out.nbytes))
        if words == 2:
            lo = out[0::2].astype(object)
            hi = out[1::2].view(np.int64).astype(object)
            return [int(h) * (1 << 64) + int(l) for l, h in zip(lo, hi)]
        return out.view(np.int64)

    def free(self):
        if self.ptr:
            self.L.vh_table_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Vec:
    """A device vector of `n` positions (the whole column is one vector)."""

    def __init__(self, n, dtype):
        self.buf = DeviceBuffer(max(n, 1) * np.dtype(dtype).itemsize)
        self.ptr = self.buf.ptr


def _count(dcount: DeviceBuffer) -> int:
    return int(dcount.to_numpy(np.int32, 1)[0])


def groupby_statements(keys, values, n):
    """HashAggr loop of relalg_translator.cpp:582-656 over device columns.

    keys: list of (vh_type, device ptr); values: list of (vh_type, device ptr)
    or None for count.  Returns (table, rows): table columns are the keys
    (64-bit) then one 128-bit column per value plus the implicit count."""
    L = _bind()
    nk, nv = len(keys), len(values)
    t = Table(nk + nv + 1, [0] * nk + [1] * (nv + 1), max(n, 1))
    hash_ = Vec(n, np.uint64)
    bucket = Vec(n, np.uint64)
    new_pos = Vec(n, np.uint64)
    eq = Vec(n, np.uint8)
    equal = Vec(n, np.uint8)
    miss, miss2, hit, hit2, found, tmp = (Vec(n, np.int32) for _ in range(6))
    cnt = DeviceBuffer(16)
    # hash = hash(k0) rehash(hash, k1) ...
    check(L.vh_vec_hash(keys[0][0], None, n, hash_.ptr, keys[0][1], None))
    for kt, kp in keys[1:]:
        check(L.vh_vec_rehash(kt, None, n, hash_.ptr, hash_.ptr, kp, None))
    # miss = <all>
    ones = DeviceBuffer.from_numpy(np.ones(max(n, 1), dtype=np.uint8))
    check(L.vh_vec_seltrue(None, n, miss.ptr, ones.ptr, cnt.ptr, None))
    miss_n = _count(cnt)
    rounds = 0
    while miss_n > 0:
        rounds += 1
        check(L.vh_vec_bucket_lookup(t.ptr, miss.ptr, miss_n, bucket.ptr, hash_.ptr, None))
        check(L.vh_vec_cmp(abi.VH_EQ, abi.VH_U64, miss.ptr, miss_n, eq.ptr, bucket.ptr, None, 0, None))
        check(L.vh_vec_selfalse(miss.ptr, miss_n, hit.ptr, eq.ptr, cnt.ptr, None))
        hit_n = _count(cnt)
        check(L.vh_vec_seltrue(miss.ptr, miss_n, miss2.ptr, eq.ptr, cnt.ptr, None))
        miss, miss2 = miss2, miss
        miss_n = _count(cnt)
        while hit_n > 0:
            # equal = check(key0) and check(key1) ...
            check(L.vh_vec_check(t.ptr, 0, hit.ptr, hit_n, equal.ptr, bucket.ptr, keys[0][0], keys[0][1], None))
            for k in range(1, nk):
                check(L.vh_vec_check(t.ptr, k, hit.ptr, hit_n, eq.ptr, bucket.ptr, keys[k][0], keys[k][1], None))
                check(L.vh_vec_arith(abi.VH_E_MUL, abi.VH_U8, hit.ptr, hit_n, equal.ptr, abi.VH_U8, equal.ptr,
                                     abi.VH_U8, eq.ptr, 0, None))     # and() on 0/1 predicates
            check(L.vh_vec_seltrue(hit.ptr, hit_n, found.ptr, equal.ptr, cnt.ptr, None))
            found_n = _count(cnt)
            for v, (vt, vp) in enumerate(values):
                if vp is None:
                    check(L.vh_vec_aggr_count(t.ptr, nk + v, found.ptr, found_n, bucket.ptr, None))
                else:
                    check(L.vh_vec_aggr_sum(t.ptr, nk + v, found.ptr, found_n, bucket.ptr, vt, vp, None))
            check(L.vh_vec_aggr_count(t.ptr, nk + nv, found.ptr, found_n, bucket.ptr, None))
            # hit = selfalse(equal); bucket = bucket_next; empty = eq(bucket, 0)
            check(L.vh_vec_selfalse(hit.ptr, hit_n, hit2.ptr, equal.ptr, cnt.ptr, None))
            h2 = _count(cnt)
            check(L.vh_vec_bucket_next(t.ptr, hit2.ptr, h2, bucket.ptr, bucket.ptr, None))
            check(L.vh_vec_cmp(abi.VH_EQ, abi.VH_U64, hit2.ptr, h2, eq.ptr, bucket.ptr, None, 0, None))
            # miss = selunion(miss, seltrue(empty) | hit): flags over positions
            check(L.vh_vec_seltrue(hit2.ptr, h2, tmp.ptr, eq.ptr, cnt.ptr, None))
            e_n = _count(cnt)
            miss_n = _selunion(L, miss, miss_n, tmp, e_n, miss2, n, cnt)
            miss, miss2 = miss2, miss
            check(L.vh_vec_selfalse(hit2.ptr, h2, hit.ptr, eq.ptr, cnt.ptr, None))
            hit_n = _count(cnt)
        # new_pos = bucket_insert(ht, hash) | miss; scatter keys | can_scatter
        check(L.vh_vec_bucket_insert(t.ptr, miss.ptr, miss_n, new_pos.ptr, hash_.ptr, None))
        check(L.vh_vec_cmp(abi.VH_EQ, abi.VH_U64, miss.ptr, miss_n, eq.ptr, new_pos.ptr, None, 0, None))
        check(L.vh_vec_selfalse(miss.ptr, miss_n, tmp.ptr, eq.ptr, cnt.ptr, None))
        can_n = _count(cnt)
        for k, (kt, kp) in enumerate(keys):
            check(L.vh_vec_scatter(t.ptr, k, tmp.ptr, can_n, new_pos.ptr, kt, kp, None))
        if rounds > 10_000:
            raise RuntimeError("group-by loop does not terminate")
    return t, t.rows()


def _selunion(L, a, an, b, bn, out, n, cnt):
    """ComplexFuncs::selunion (runtime.cpp:67-169) as flags over positions:
    clear a u8 vector, mark both lists, compact it (ascending, no
    duplicates)."""
    flags = Vec(n, np.uint8)
    one = DeviceBuffer.from_numpy(np.ones(max(n, 1), dtype=np.uint8))
    # flags[i] = (1 != 1) = 0 everywhere, then 1 at the positions of a and b
    check(L.vh_vec_cmp(abi.VH_NE, abi.VH_U8, None, n, flags.ptr, one.ptr, None, 1, None))
    for s, sn in ((a, an), (b, bn)):
        check(L.vh_vec_cmp(abi.VH_EQ, abi.VH_U8, s.ptr, sn, flags.ptr, one.ptr, None, 1, None))
    check(L.vh_vec_seltrue(None, n, out.ptr, flags.ptr, cnt.ptr, None))
    return _count(cnt)


def join_statements(build_key, build_pays, nb, probe_key, n, variant):
    """HashJoin of relalg_translator.cpp:828-1111 as statements.

    Build (pipelines A+B, :828-928): hash the build keys, write_pos + write
    the key, payload and hash columns, bucket_build.  Probe (pipeline C,
    :933-1111): bucket_lookup, then while tuples are active: check the key,
    gather the payloads of the matches, drop the matched tuples when the join
    is Join01 (first match only), bucket_next, keep the non-empty buckets.

    build_key / probe_key: (vh_type, device ptr); build_pays: list of
    (vh_type, device ptr).  Returns (positions, [payload arrays]): one entry
    per (probe position, matching build row), int64 payloads."""
    L = _bind()
    npay = len(build_pays)
    t = Table(1 + npay, [0] * (1 + npay), max(nb, 1))
    bhash = Vec(nb, np.uint64)
    cnt = DeviceBuffer(16)
    if nb:
        check(L.vh_vec_hash(build_key[0], None, nb, bhash.ptr, build_key[1], None))
        base = C.c_int64()
        check(L.vh_vec_write_pos(t.ptr, nb, C.byref(base)))
        check(L.vh_vec_write(t.ptr, 0, base.value, None, nb, build_key[0], build_key[1], None))
        for j, (pt, pp) in enumerate(build_pays):
            check(L.vh_vec_write(t.ptr, 1 + j, base.value, None, nb, pt, pp, None))
        check(L.vh_vec_write(t.ptr, -1, base.value, None, nb, abi.VH_U64, bhash.ptr, None))
    check(L.vh_vec_bucket_build(t.ptr, None))

    phash = Vec(n, np.uint64)
    bucket = Vec(n, np.uint64)
    eq, nz = Vec(n, np.uint8), Vec(n, np.uint8)
    active, active2, match = (Vec(n, np.int32) for _ in range(3))
    gathered = [Vec(n, np.int64) for _ in range(npay)]
    positions, pays = [], [[] for _ in range(npay)]
    check(L.vh_vec_hash(probe_key[0], None, n, phash.ptr, probe_key[1], None))
    check(L.vh_vec_bucket_lookup(t.ptr, None, n, bucket.ptr, phash.ptr, None))
    check(L.vh_vec_cmp(abi.VH_NE, abi.VH_U64, None, n, nz.ptr, bucket.ptr, None, 0, None))
    check(L.vh_vec_seltrue(None, n, active.ptr, nz.ptr, cnt.ptr, None))
    act_n = _count(cnt)
    rounds = 0
    while act_n > 0:
        rounds += 1
        check(L.vh_vec_check(t.ptr, 0, active.ptr, act_n, eq.ptr, bucket.ptr, probe_key[0], probe_key[1], None))
        check(L.vh_vec_seltrue(active.ptr, act_n, match.ptr, eq.ptr, cnt.ptr, None))
        m_n = _count(cnt)
        if m_n:
            for j in range(npay):
                check(L.vh_vec_gather(t.ptr, 1 + j, match.ptr, m_n, gathered[j].ptr, abi.VH_I64, bucket.ptr, None))
            pos = match.buf.to_numpy(np.int32, m_n)
            positions.append(pos)
            for j in range(npay):
                pays[j].append(gathered[j].buf.to_numpy(np.int64, n)[pos])
        if variant == abi.VH_JOIN01:
            # matched tuples are done: active = selfalse(eq) | active
            check(L.vh_vec_selfalse(active.ptr, act_n, active2.ptr, eq.ptr, cnt.ptr, None))
            active, active2 = active2, active
            act_n = _count(cnt)
        check(L.vh_vec_bucket_next(t.ptr, active.ptr, act_n, bucket.ptr, bucket.ptr, None))
        check(L.vh_vec_cmp(abi.VH_NE, abi.VH_U64, active.ptr, act_n, nz.ptr, bucket.ptr, None, 0, None))
        check(L.vh_vec_seltrue(active.ptr, act_n, active2.ptr, nz.ptr, cnt.ptr, None))
        active, active2 = active2, active
        act_n = _count(cnt)
        if rounds > max(nb, 1) + 1:
            raise RuntimeError("join probe loop does not terminate")
    t.free()
    if not positions:
        return np.zeros(0, np.int32), [np.zeros(0, np.int64) for _ in range(npay)]
    return np.concatenate(positions), [np.concatenate(p) for p in pays]
