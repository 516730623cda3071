"""`.tbl` import (CPU): the casts of the reference's importer, its MinMaxInfo
bookkeeping, the binary column cache, and Q1 / Q6 over imported columns
through the oracle against the same queries computed from the text."""
import datetime
import os

import numpy as np
import pytest

import oracle
from vldb_voila_amd import tbl

from tbl_data import Q1_COLS, q1, q1_from_text, q6, q6_from_text, write_lineitem


def test_integer_cast():
    assert tbl.cast_integer("42") == 42
    assert tbl.cast_integer("  -7 ") == -7
    assert tbl.cast_integer("+3") == 3
    assert tbl.cast_integer("12.9") == 12          # digits after '.' are ignored
    for bad in ("", "-", "1a", "12345678901"):       # 11 digits: more than a 32-bit integer's 10
        with pytest.raises(tbl.FormatError):
            tbl.cast_integer(bad)


def test_numeric_cast_scales_to_the_precision():
    assert tbl.cast_numeric("12.5", 2) == 1250
    assert tbl.cast_numeric("-0.07", 2) == -7
    assert tbl.cast_numeric("3", 2) == 300
    assert tbl.cast_numeric("0.10", 2) == 10
    assert tbl.cast_numeric("1.2300", 2) == 123      # trailing zeros of the fraction do not count
    assert tbl.cast_numeric("104949.50", 2) == 10494950
    for bad in ("1.234", "1.2.3", "1e3", "", "x"):
        with pytest.raises(tbl.FormatError):
            tbl.cast_numeric(bad, 2)


def test_julian_day_matches_the_proleptic_gregorian_ordinal():
    # JDN = ordinal (days since 0001-01-01, day 1) + 1721425, for every date
    day = datetime.date(1900, 1, 1)
    while day.year < 2101:
        assert tbl.julian_day(day.year, day.month, day.day) == day.toordinal() + 1721425
        day += datetime.timedelta(days=37)
    assert tbl.cast_date("1970-01-01") == 2440588
    assert tbl.cast_date("1998-09-02") == 2451059
    for bad in ("1998-13-01", "1998-00-10", "1998-1-32", "19980902", "1998-09-"):
        with pytest.raises(tbl.FormatError):
            tbl.cast_date(bad)


def test_char1_is_its_byte():
    assert tbl.cast_char1("R") == 82 and tbl.cast_char1("N") == 78


def test_minmax_keeps_the_reference_initial_maximum():
    assert tbl.minmax(np.array([-5, -2], dtype=np.int64)) == (-5, 0)   # hi never drops below DBL_MIN
    assert tbl.minmax(np.array([3, 9], dtype=np.int32)) == (3, 9)
    assert tbl.minmax(np.array([], dtype=np.int32)) == (0, 0)


def test_load_writes_the_cache_and_reads_it_back(tmp_path):
    write_lineitem(tmp_path / "lineitem.tbl", 500, seed=3)
    a, b = tbl.load_table(str(tmp_path), "lineitem", Q1_COLS)
    for c in Q1_COLS:
        assert os.path.exists(tmp_path / "cached" / f"lineitem_{c}")
    assert a["l_returnflag"].dtype == np.uint8 and a["l_extendedprice"].dtype == np.int64
    assert set(np.unique(a["l_returnflag"])) <= {ord("R"), ord("A"), ord("N")}
    assert b["l_discount"] == (int(a["l_discount"].min()), int(a["l_discount"].max()))
    os.remove(tmp_path / "lineitem.tbl")             # later loads need only the cache
    a2, _ = tbl.load_table(str(tmp_path), "lineitem", Q1_COLS)
    for c in Q1_COLS:
        assert np.array_equal(a[c], a2[c])
    with pytest.raises(ValueError):
        tbl.load_table(str(tmp_path), "lineitem", ["l_nosuchcolumn"])


def test_load_char_and_varchar_columns(tmp_path):
    """char(n) / varchar(n) columns load as (1 + n)-byte slots, in the cache
    too, and read back the text's fields (char: leading spaces dropped)."""
    from vldb_voila_amd import varchar
    write_lineitem(tmp_path / "lineitem.tbl", 200, seed=4)
    cols = ["l_shipinstruct", "l_shipmode", "l_comment"]
    a, b = tbl.load_table(str(tmp_path), "lineitem", cols)
    assert a["l_shipinstruct"].shape == (200, 26) and a["l_comment"].shape == (200, 45)
    assert b["l_comment"] is None
    assert os.path.getsize(tmp_path / "cached" / "lineitem_l_comment") == 200 * 45
    fields = [line.split("|") for line in open(tmp_path / "lineitem.tbl")]
    assert varchar.unpack(a["l_shipinstruct"]) == [f[13].lstrip(" ").encode() for f in fields]
    assert varchar.unpack(a["l_comment"]) == [f[15].encode() for f in fields]
    a2, _ = tbl.load_table(str(tmp_path), "lineitem", cols)
    assert all(np.array_equal(a[c], a2[c]) for c in cols)


def test_q1_q6_on_imported_columns_match_the_text(tmp_path):
    path = tmp_path / "lineitem.tbl"
    write_lineitem(path, 3000, seed=5)
    a, b = tbl.load_table(str(tmp_path), "lineitem", Q1_COLS)
    got = oracle.sorted_rows(oracle.bruteforce(q1(a, b).desc()))
    assert got == q1_from_text(path)
    s, n = q6_from_text(path)
    assert n > 0
    assert oracle.sorted_rows(oracle.bruteforce(q6(a, b).desc())) == [str(s)]


def test_relalg_q1_q6_plans_over_imported_columns(tmp_path):
    """The same queries written as relalg plans over the imported catalog
    lower to pipelines with the same answers."""
    from vldb_voila_amd import relalg
    from tbl_data import Q1_CUTOFF, Q6_FROM, Q6_TO
    path = tmp_path / "lineitem.tbl"
    write_lineitem(path, 2000, seed=9)
    a, b = tbl.load_table(str(tmp_path), "lineitem", Q1_COLS)
    db = relalg.Database().add_table("lineitem", a, bounds=b)
    assert oracle.sorted_rows(oracle.bruteforce(db.lower(relalg.tpch_q1_plan(Q1_CUTOFF)))) == q1_from_text(path)
    d6 = db.lower(relalg.tpch_q6_plan(Q6_FROM, Q6_TO))
    assert d6.probe.npreds == 5
    assert oracle.sorted_rows(oracle.bruteforce(d6)) == [str(q6_from_text(path)[0])]
