"""Varchar columns and their statements on the GPU.

A varchar column holds one (1 + max_len)-byte slot per row: a length byte,
then the bytes, zero-padded.  This is the binary column of the reference's
Char<n> / Varchar<n> (db-engine-paradigms Types.hpp:125-130, :243-250) that
varchar(void* data, idx, max_len) reads (runtime.hpp:192-205).

The statements are the vec_* primitives the reference generates for varchar
operands (kernel_generator.py:412-468): hash / rehash (MurmurHash64A,
runtime.hpp:434-484), contains (memmem, runtime.hpp:206-208) and eq / ne
(varchar::operator==, runtime.hpp:154-156).  Their u8 results serve as
predicate columns of a pipeline: `select(name, "==", 1)`.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import abi
from .runtime import DeviceBuffer, check, lib

MAX_LEN = 255


def _bind():
    L = lib()
    vp, i32, i64 = C.c_void_p, C.c_int32, C.c_int64
    L.vh_vec_hash_varchar.argtypes = [vp, i64, vp, vp, i32, vp]
    L.vh_vec_rehash_varchar.argtypes = [vp, i64, vp, vp, vp, i32, vp]
    L.vh_vec_contains.argtypes = [vp, i64, vp, vp, i32, C.c_char_p, i32, vp]
    L.vh_vec_cmp_varchar.argtypes = [i32, vp, i64, vp, vp, i32, vp, i32, C.c_char_p, i32, vp]
    return L


def _bytes(v) -> bytes:
    return v.encode("latin-1") if isinstance(v, str) else bytes(v)


def pack(values, max_len: int) -> np.ndarray:
    """Host column of (1 + max_len)-byte slots, shape (rows, 1 + max_len)."""
    if not 1 <= max_len <= MAX_LEN:
        raise ValueError(f"max_len {max_len} outside 1..{MAX_LEN}")
    out = np.zeros((len(values), max_len + 1), dtype=np.uint8)
    for i, v in enumerate(values):
        b = _bytes(v)
        if len(b) > max_len:
            raise ValueError(f"row {i}: {len(b)} bytes exceed max_len {max_len}")
        out[i, 0] = len(b)
        out[i, 1:1 + len(b)] = np.frombuffer(b, dtype=np.uint8)
    return out


def unpack(col: np.ndarray) -> list:
    """The rows of a slot column as bytes."""
    return [bytes(r[1:1 + min(int(r[0]), col.shape[1] - 1)]) for r in col]


def _sel(sel):
    return sel.ptr if sel is not None else None


def hash(col_ptr: int, rows: int, max_len: int, prev: DeviceBuffer | None = None, sel=None,
         out: DeviceBuffer | None = None, stream=None) -> DeviceBuffer:
    """vec_hash (prev None) / vec_rehash on a device varchar column: u64 per row."""
    L = _bind()
    out = out or DeviceBuffer(max(rows, 1) * 8)
    n = rows if sel is None else sel.count
    s = stream.ptr if stream else None
    if prev is None:
        check(L.vh_vec_hash_varchar(_sel(sel), n, out.ptr, col_ptr, max_len, s))
    else:
        check(L.vh_vec_rehash_varchar(_sel(sel), n, out.ptr, prev.ptr, col_ptr, max_len, s))
    return out


def contains(col_ptr: int, rows: int, max_len: int, needle, sel=None, out: DeviceBuffer | None = None,
             stream=None) -> DeviceBuffer:
    """vec_contains: u8 per row, 1 where the row contains `needle`."""
    L = _bind()
    nb = _bytes(needle)
    out = out or DeviceBuffer(max(rows, 1))
    n = rows if sel is None else sel.count
    check(L.vh_vec_contains(_sel(sel), n, out.ptr, col_ptr, max_len, nb, len(nb), stream.ptr if stream else None))
    return out


def compare(cmp: str, col_ptr: int, rows: int, max_len: int, other, other_max_len: int = 0, sel=None,
            out: DeviceBuffer | None = None, stream=None) -> DeviceBuffer:
    """vec_eq / vec_ne: u8 per row against a constant (str / bytes) or a
    second device column (its address, with other_max_len)."""
    L = _bind()
    op = abi.CMP_NAMES[cmp]
    if op not in (abi.VH_EQ, abi.VH_NE):
        raise ValueError("varchar columns compare with == and != only")
    out = out or DeviceBuffer(max(rows, 1))
    n = rows if sel is None else sel.count
    s = stream.ptr if stream else None
    if isinstance(other, (str, bytes, bytearray)):
        cb = _bytes(other)
        check(L.vh_vec_cmp_varchar(op, _sel(sel), n, out.ptr, col_ptr, max_len, None, 0, cb, len(cb), s))
    else:
        check(L.vh_vec_cmp_varchar(op, _sel(sel), n, out.ptr, col_ptr, max_len, int(other), other_max_len,
                                   None, 0, s))
    return out


class Selection:
    """A device selection vector (ascending int32 positions) for the calls above."""

    def __init__(self, positions: np.ndarray):
        pos = np.ascontiguousarray(positions, dtype=np.int32)
        self.count = len(pos)
        self.buf = DeviceBuffer.from_numpy(pos) if len(pos) else DeviceBuffer(4)
        self.ptr = self.buf.ptr
