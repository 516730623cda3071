"""CPU oracle loader.  TEST INFRASTRUCTURE: only tests/, __graft_entry__.smoke()
and bench.py's cpu_baseline leg import this package, as the checker."""
from __future__ import annotations

import ctypes as C
import os
import subprocess

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "libvoila_oracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        L.vo_run_pipeline.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_char_p,
                                      C.c_size_t, C.POINTER(C.c_size_t),
                                      C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
        L.vo_bruteforce_pipeline.argtypes = [C.c_void_p, C.c_char_p, C.c_size_t,
                                             C.POINTER(C.c_size_t), C.POINTER(C.c_int64)]
        L.vo_hash.restype = C.c_uint64
        L.vo_hash.argtypes = [C.c_int64]
        L.vo_rehash.restype = C.c_uint64
        L.vo_rehash.argtypes = [C.c_uint64, C.c_int64]
        L.vo_calc_num_buckets.restype = C.c_size_t
        L.vo_calc_num_buckets.argtypes = [C.c_size_t, C.c_size_t, C.c_size_t]
        L.vo_to_cstr_i128.restype = C.c_size_t
        L.vo_to_cstr_i128.argtypes = [C.c_uint64, C.c_uint64, C.c_char_p]
        L.vo_seltrue.restype = C.c_int32
        L.vo_seltrue.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p]
        L.vo_selfalse.restype = C.c_int32
        L.vo_selfalse.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p]
        L.vo_selunion.restype = C.c_int32
        L.vo_selunion.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p, C.c_int32]
        L.vo_type_from_minmax.restype = C.c_int
        L.vo_type_from_minmax.argtypes = [C.c_double, C.c_double]
        L.vo_typing_of.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.vo_sizeof_pipeline_desc.restype = C.c_size_t
        L.vo_murmur64a.restype = C.c_uint64
        L.vo_murmur64a.argtypes = [C.c_char_p, C.c_size_t, C.c_uint64]
        L.vo_vec_hash_varchar.argtypes = [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]
        L.vo_vec_contains.argtypes = [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_int, C.c_char_p,
                                      C.c_int]
        L.vo_vec_cmp_varchar.argtypes = [C.c_int, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_int,
                                         C.c_void_p, C.c_int, C.c_char_p, C.c_int]
        _LIB = L
    return _LIB


def run_pipeline(desc, threads=1, vector_size=1024, morsel_size=16 * 1024, out_cap=1 << 20):
    """vector(1024) restatement; returns (text rows, elapsed_ns of the run)."""
    L = lib()
    n = C.c_size_t(0)
    rows = C.c_int64(0)
    ns = C.c_int64(0)
    buf = C.create_string_buffer(out_cap)
    rc = L.vo_run_pipeline(C.byref(desc), threads, vector_size, morsel_size, buf, out_cap,
                           C.byref(n), C.byref(rows), C.byref(ns))
    if rc == 5:   # output larger than the buffer: run again with the right size
        buf = C.create_string_buffer(n.value + 1)
        rc = L.vo_run_pipeline(C.byref(desc), threads, vector_size, morsel_size, buf,
                               n.value + 1, C.byref(n), C.byref(rows), C.byref(ns))
    if rc:
        raise RuntimeError(f"vo_run_pipeline failed: {rc}")
    return buf.value.decode(), ns.value


def bruteforce(desc):
    L = lib()
    n = C.c_size_t(0)
    rows = C.c_int64(0)
    rc = L.vo_bruteforce_pipeline(C.byref(desc), None, 0, C.byref(n), C.byref(rows))
    if rc not in (0, 5):
        raise RuntimeError(f"vo_bruteforce_pipeline failed: {rc}")
    buf = C.create_string_buffer(n.value + 1)
    rc = L.vo_bruteforce_pipeline(C.byref(desc), buf, n.value + 1, C.byref(n), C.byref(rows))
    if rc:
        raise RuntimeError(f"vo_bruteforce_pipeline failed: {rc}")
    return buf.value.decode()


def sorted_rows(text: str):
    """Query::check_result compares line-split, sorted rows (runtime_framework.cpp:219-279)."""
    return sorted(l for l in text.split("\n") if l)
