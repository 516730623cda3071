"""TPC-H style `.tbl` import with a binary column cache.

The reference's benchmarks load their tables through db-engine-paradigms'
importer: every `|`-terminated field of `<dir>/<table>.tbl` is cast to its
column's type (Import.cpp:165-247, the type casts in runtime/Types.cpp and
runtime/Types.hpp), each column is written once to `<dir>/cached/<table>_<col>`
as raw values and read back from there on later loads, and its [min, max] is
collected for the typing pass (MinMaxInfo, runtime/Types.hpp:11-33; the
typing's use at Import.cpp:175-186).  This module does the same for the
column types the operator path handles:

  integer  Integer::castString (runtime/Types.cpp:54-95) -> int32
  bigint   the same digits -> int64
  numeric  Numeric<len, p>::castString (runtime/Types.hpp:628-683): value
           scaled by 10^p, trailing zeros of the fraction dropped -> int64
  date     Date::castString: YYYY-MM-DD -> Julian day number
           (mergeJulianDay, runtime/Types.cpp:108-117) -> int32
  char1    Char<1>::castString: the first byte -> uint8
  char(n)  Char<n>::castString (Types.hpp:295-305): leading spaces dropped,
           at most n bytes -> a (1 + n)-byte slot: length byte, then bytes
  varchar(n) Varchar<n>::castString (Types.hpp:165-171): the bytes as they
           are, at most n -> the same slot layout

Char and varchar columns load as uint8 arrays of shape (rows, 1 + n), the
layout of the reference's binary Char<n> / Varchar<n> columns (a u8 length
indicator, then `char value[n]`; Types.hpp:125-130, :243-250), with the
bytes after the length zeroed.  The string statements of the path
(vldb_voila_amd.varchar) read them on the GPU.  Column types follow the
reference's TPC-H schema (Import.cpp:262-392).
"""
from __future__ import annotations

import os
import sys

import numpy as np

DBL_MAX = sys.float_info.max
DBL_MIN = sys.float_info.min   # MinMaxInfo starts hi at numeric_limits<double>::min()

NUMPY_TYPE = {"integer": np.int32, "bigint": np.int64, "numeric": np.int64, "date": np.int32,
              "char1": np.uint8, "char": np.uint8, "varchar": np.uint8}
STRING_TYPES = ("char", "varchar")

# (column, type, numeric precision or string length) of the reference's
# TPC-H tables
TPCH = {
    "part": [("p_partkey", "integer", 0), ("p_name", "varchar", 55), ("p_mfgr", "char", 25),
             ("p_brand", "char", 10), ("p_type", "varchar", 25), ("p_size", "integer", 0),
             ("p_container", "char", 10), ("p_retailprice", "numeric", 2), ("p_comment", "varchar", 23)],
    "supplier": [("s_suppkey", "integer", 0), ("s_name", "char", 25), ("s_address", "varchar", 40),
                 ("s_nationkey", "integer", 0), ("s_phone", "char", 15), ("s_acctbal", "numeric", 2),
                 ("s_comment", "varchar", 101)],
    "partsupp": [("ps_partkey", "integer", 0), ("ps_suppkey", "integer", 0), ("ps_availqty", "integer", 0),
                 ("ps_supplycost", "numeric", 2), ("ps_comment", "varchar", 199)],
    "customer": [("c_custkey", "integer", 0), ("c_name", "char", 25), ("c_address", "varchar", 40),
                 ("c_nationkey", "integer", 0), ("c_phone", "char", 15), ("c_acctbal", "numeric", 2),
                 ("c_mktsegment", "char", 10), ("c_comment", "varchar", 117)],
    "orders": [("o_orderkey", "integer", 0), ("o_custkey", "integer", 0), ("o_orderstatus", "char1", 0),
               ("o_totalprice", "numeric", 2), ("o_orderdate", "date", 0), ("o_orderpriority", "char", 15),
               ("o_clerk", "char", 15), ("o_shippriority", "integer", 0), ("o_comment", "varchar", 79)],
    "lineitem": [("l_orderkey", "integer", 0), ("l_partkey", "integer", 0), ("l_suppkey", "integer", 0),
                 ("l_linenumber", "integer", 0), ("l_quantity", "numeric", 2),
                 ("l_extendedprice", "numeric", 2), ("l_discount", "numeric", 2), ("l_tax", "numeric", 2),
                 ("l_returnflag", "char1", 0), ("l_linestatus", "char1", 0), ("l_shipdate", "date", 0),
                 ("l_commitdate", "date", 0), ("l_receiptdate", "date", 0), ("l_shipinstruct", "char", 25),
                 ("l_shipmode", "char", 10), ("l_comment", "varchar", 44)],
    "nation": [("n_nationkey", "integer", 0), ("n_name", "char", 25), ("n_regionkey", "integer", 0),
               ("n_comment", "varchar", 152)],
    "region": [("r_regionkey", "integer", 0), ("r_name", "char", 25), ("r_comment", "varchar", 152)],
}


class FormatError(ValueError):
    pass


def _trim(s: str) -> str:
    return s.strip(" ")


def _sign(s: str):
    if s[:1] == "-":
        return True, s[1:]
    if s[:1] == "+":
        return False, s[1:]
    return False, s


def cast_integer(s: str) -> int:
    """Integer::castString: optional sign, decimal digits, anything after a
    '.' ignored, at most 10 digits."""
    neg, s = _sign(_trim(s))
    if not s:
        raise FormatError("invalid number format: found non-integer characters")
    v = 0
    digits = 0
    for c in s:
        if "0" <= c <= "9":
            v = v * 10 + (ord(c) - 48)
            digits += 1
        elif c == ".":
            break
        else:
            raise FormatError("invalid number format: invalid character in integer string")
    if digits > 10:
        raise FormatError("invalid number format: too many characters")
    return -v if neg else v


def cast_bigint(s: str) -> int:
    neg, s = _sign(_trim(s))
    if not s or not s.isdigit():
        raise FormatError("invalid number format")
    return -int(s) if neg else int(s)


def cast_numeric(s: str, precision: int) -> int:
    """Numeric<len, precision>::castString: the digits before and after the
    point as one integer, the fraction's trailing zeros dropped, then scaled
    to `precision` fraction digits; more fraction digits than that, or more
    than 18 integer digits, is an error."""
    neg, s = _sign(_trim(s))
    if not s:
        raise FormatError("invalid number format: found non-numeric characters")
    if "." in s:
        ip, fp = s.split(".", 1)
        if "." in fp:
            raise FormatError("invalid number format: already in fraction")
        fp = fp.rstrip("0")
    else:
        ip, fp = s, ""
    if (ip and not ip.isdigit()) or (fp and not fp.isdigit()):
        raise FormatError("invalid number format: invalid character in numeric string")
    if len(ip) > 18 or len(fp) > precision:
        raise FormatError("invalid number format: loosing precision")
    v = int(ip + fp) if (ip + fp) else 0
    v *= 10 ** (precision - len(fp))
    return -v if neg else v


def julian_day(year: int, month: int, day: int) -> int:
    """mergeJulianDay (the Calendar FAQ's algorithm), in unsigned arithmetic."""
    a = (14 - month) // 12
    y = year + 4800 - a
    m = month + 12 * a - 3
    return day + (153 * m + 2) // 5 + 365 * y + y // 4 - y // 100 + y // 400 - 32045


def cast_date(s: str) -> int:
    """Date::castString: YYYY-MM-DD (digits only, '-' separated)."""
    s = _trim(s)
    parts = s.split("-")
    if len(parts) != 3 or not all(p.isdigit() for p in parts[:2]) or (parts[2] and not parts[2].isdigit()):
        raise FormatError("invalid date format")
    year, month, day = int(parts[0]), int(parts[1]), int(parts[2] or 0)
    if year > 9999 or not 1 <= month <= 12 or not 1 <= day <= 31:
        raise FormatError("invalid date format")
    return julian_day(year, month, day)


def cast_char1(s: str) -> int:
    """Char<1>::castString: the first byte."""
    b = s.encode("latin-1")
    return b[0] if b else 0


def cast_char(s: str, n: int) -> bytes:
    """Char<n>::castString: leading spaces dropped, at most n bytes."""
    b = s.encode("latin-1").lstrip(b" ")
    if len(b) > n:
        raise FormatError(f"char({n}) value of {len(b)} bytes")
    return b


def cast_varchar(s: str, n: int) -> bytes:
    """Varchar<n>::castString: the bytes as they are, at most n."""
    b = s.encode("latin-1")
    if len(b) > n:
        raise FormatError(f"varchar({n}) value of {len(b)} bytes")
    return b


def cast(s: str, kind: str, precision: int = 0):
    if kind == "integer":
        return cast_integer(s)
    if kind == "bigint":
        return cast_bigint(s)
    if kind == "numeric":
        return cast_numeric(s, precision)
    if kind == "date":
        return cast_date(s)
    if kind == "char1":
        return cast_char1(s)
    if kind == "char":
        return cast_char(s, precision)
    if kind == "varchar":
        return cast_varchar(s, precision)
    raise ValueError(f"column type {kind!r} is not loadable")


def minmax(values: np.ndarray):
    """MinMaxInfo over a column: lo starts at DBL_MAX and hi at DBL_MIN, the
    smallest positive double (numeric_limits<double>::min(), as the
    reference initialises it), so a column with no value above zero reports
    hi = DBL_MIN, which the integer column metadata holds as 0.  The bounds
    are kept exact (the reference's doubles round integers beyond 2^53; the
    kernels check keys against these bounds, so they must not be narrower
    than the data)."""
    if not len(values):
        return 0, 0
    lo = int(values.min())
    hi = int(values.max())
    return lo, (hi if hi >= 1 else 0)


def _slots(values, n: int) -> np.ndarray:
    """(1 + n)-byte slots: the length byte, then the bytes, zero-padded."""
    out = np.zeros((len(values), n + 1), dtype=np.uint8)
    for i, b in enumerate(values):
        out[i, 0] = len(b)
        out[i, 1:1 + len(b)] = np.frombuffer(b, dtype=np.uint8)
    return out


def _cache_path(cachedir, table, col):
    return os.path.join(cachedir, f"{table}_{col}")


def load_table(directory: str, table: str, columns=None, schema=None):
    """Columns of `<directory>/<table>.tbl` as numpy arrays, with their
    [min, max] (None for char / varchar columns).  The first load writes
    `<directory>/cached/<table>_<col>` (raw values; string columns as their
    slots); later loads read those, as the reference's importer does.
    `columns`: names to load (default: every column of the schema);
    `schema`: [(name, type, precision or length)] for tables outside TPC-H."""
    spec = schema or TPCH[table]
    kinds = {name: (kind, prec) for name, kind, prec in spec}
    want = columns or [name for name, kind, _ in spec if kind]
    for c in want:
        if c not in kinds or kinds[c][0] is None:
            raise ValueError(f"{table}.{c} is not a loadable column")
    cachedir = os.path.join(directory, "cached")
    os.makedirs(cachedir, exist_ok=True)
    if not all(os.path.exists(_cache_path(cachedir, table, c)) for c in want):
        idx = {name: i for i, (name, _, _) in enumerate(spec)}
        vals = {c: [] for c in want}
        with open(os.path.join(directory, f"{table}.tbl"), encoding="latin-1") as f:
            for line in f:
                fields = line.rstrip("\n").split("|")
                for c in want:
                    kind, prec = kinds[c]
                    vals[c].append(cast(fields[idx[c]], kind, prec))
        for c in want:
            kind, n = kinds[c]
            if kind in STRING_TYPES:
                _slots(vals[c], n).tofile(_cache_path(cachedir, table, c))
            else:
                np.asarray(vals[c], dtype=NUMPY_TYPE[kind]).tofile(_cache_path(cachedir, table, c))
    out, bounds = {}, {}
    sizes = set()
    for c in want:
        kind, n = kinds[c]
        a = np.fromfile(_cache_path(cachedir, table, c), dtype=NUMPY_TYPE[kind])
        if kind in STRING_TYPES:
            a = a.reshape(-1, n + 1)
            bounds[c] = None
        else:
            bounds[c] = minmax(a)
        out[c] = a
        sizes.add(len(a))
    if len(sizes) > 1:
        raise ValueError(f"columns of {table} differ in size")
    return out, bounds
