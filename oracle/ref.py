"""Loader for oracle/_ref/libvoila_ref.so: the reference's own runtime.cpp /
runtime.hpp compiled from /root/reference (oracle/Makefile, target `ref`).
TEST INFRASTRUCTURE: used only to pin the oracle."""
from __future__ import annotations

import ctypes as C
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
PATH = os.path.join(HERE, "_ref", "libvoila_ref.so")
REFERENCE = "/root/reference"
_LIB = None


def available() -> bool:
    if os.path.exists(PATH):
        return True
    if os.path.exists(os.path.join(REFERENCE, "runtime.cpp")):
        r = subprocess.run(["make", "-s", "-C", HERE, "ref"], capture_output=True)
        return r.returncode == 0 and os.path.exists(PATH)
    return False


def lib():
    global _LIB
    if _LIB is None:
        L = C.CDLL(PATH)
        u64, i64, i32, u32 = C.c_uint64, C.c_int64, C.c_int32, C.c_uint32
        for name, arg in (("ref_hash_i32", i32), ("ref_hash_u32", u32), ("ref_hash_i64", i64),
                          ("ref_hash_u64", u64), ("ref_hash_u8", C.c_uint8)):
            getattr(L, name).restype = u64
            getattr(L, name).argtypes = [arg]
        for name, arg in (("ref_rehash_i32", i32), ("ref_rehash_u32", u32), ("ref_rehash_i64", i64)):
            getattr(L, name).restype = u64
            getattr(L, name).argtypes = [u64, arg]
        L.ref_to_cstr_i128.restype = C.c_size_t
        L.ref_to_cstr_i128.argtypes = [u64, u64, C.c_char_p]
        L.ref_to_cstr_i64.restype = C.c_size_t
        L.ref_to_cstr_i64.argtypes = [i64, C.c_char_p]
        L.ref_to_cstr_u64.restype = C.c_size_t
        L.ref_to_cstr_u64.argtypes = [u64, C.c_char_p]
        for name in ("ref_select_true", "ref_select_false"):
            getattr(L, name).restype = i32
            getattr(L, name).argtypes = [C.c_void_p, i32, C.c_void_p, C.c_void_p]
        L.ref_selunion.restype = i32
        L.ref_selunion.argtypes = [C.c_void_p, C.c_void_p, i32, C.c_void_p, i32]
        L.ref_avx512_hash_u32.restype = i32
        L.ref_avx512_hash_u32.argtypes = [C.c_void_p, i32, C.c_void_p, C.c_void_p]
        L.ref_bucket_lookup.restype = i32
        L.ref_bucket_lookup.argtypes = [C.c_void_p, i32, C.c_void_p, C.c_void_p, C.c_void_p, u64]
        L.ref_bucket_next.restype = i32
        L.ref_bucket_next.argtypes = [C.c_void_p, i32, C.c_void_p, C.c_void_p, u64]
        L.ref_check_u32.restype = i32
        L.ref_check_u32.argtypes = [C.c_void_p, i32, C.c_void_p, C.c_void_p, C.c_void_p, u64]
        L.ref_hash_varchar.restype = u64
        L.ref_hash_varchar.argtypes = [C.c_void_p, C.c_size_t, C.c_size_t]
        L.ref_rehash_varchar.restype = u64
        L.ref_rehash_varchar.argtypes = [u64, C.c_void_p, C.c_size_t, C.c_size_t]
        L.ref_varchar_contains.restype = C.c_int
        L.ref_varchar_contains.argtypes = [C.c_void_p, C.c_size_t, C.c_size_t, C.c_char_p, C.c_uint32]
        L.ref_varchar_eq.restype = C.c_int
        L.ref_varchar_eq.argtypes = [C.c_void_p, C.c_size_t, C.c_size_t, C.c_void_p, C.c_size_t, C.c_size_t]
        _LIB = L
    return _LIB
