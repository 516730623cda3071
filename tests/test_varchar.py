"""Varchar statements of the oracle (CPU), pinned three ways: MurmurHash64A
against SMHasher's published verification value for it (0x1F0D3804,
"Murmur2B"); hash, rehash, contains and equality against golden vectors the
reference's own runtime.hpp wrote (tests/golden/make_varchar_vectors.py) and,
when the reference sources are present, against that runtime live; contains
and equality against Python's own bytes routines.  Also the slot layout and
the importer's char / varchar casts."""
import json
import os

import numpy as np
import pytest

import oracle
from oracle import ref
from vldb_voila_amd import abi, tbl, varchar

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "varchar_vectors.json")


@pytest.fixture(scope="module")
def golden():
    with open(GOLDEN) as f:
        g = json.load(f)
    g["values"] = [bytes.fromhex(s) for s in g["strings"]]
    return g


def oracle_hash(col, max_len, prev=None):
    L = oracle.lib()
    res = np.zeros(len(col), dtype=np.uint64)
    L.vo_vec_hash_varchar(None, len(col), res.ctypes.data, prev.ctypes.data if prev is not None else None,
                          col.ctypes.data, max_len)
    return res


def oracle_contains(col, max_len, needle):
    res = np.zeros(len(col), dtype=np.uint8)
    oracle.lib().vo_vec_contains(None, len(col), res.ctypes.data, col.ctypes.data, max_len, needle, len(needle))
    return res


def oracle_eq(col1, m1, col2, m2):
    res = np.zeros(len(col1), dtype=np.uint8)
    oracle.lib().vo_vec_cmp_varchar(abi.VH_EQ, None, len(col1), res.ctypes.data, col1.ctypes.data, m1,
                                    col2.ctypes.data, m2, None, 0)
    return res


def test_murmur64a_smhasher_verification_value():
    """SMHasher's VerificationTest: hash keys {}, {0}, {0,1}, ... of 0..255
    bytes with seed 256 - n, hash the concatenated 64-bit results with seed
    0; the low 32 bits are 0x1F0D3804 for MurmurHash64A."""
    L = oracle.lib()
    key = bytes(range(256))
    hashes = b"".join(L.vo_murmur64a(key[:n], n, 256 - n).to_bytes(8, "little") for n in range(256))
    final = L.vo_murmur64a(hashes, len(hashes), 0)
    assert final & 0xFFFFFFFF == 0x1F0D3804


def test_hash_and_rehash_match_reference_vectors(golden):
    m = golden["max_len"]
    col = varchar.pack(golden["values"], m)
    assert [str(int(h)) for h in oracle_hash(col, m)] == golden["hash"]
    seeds = np.array([int(s) for s, _ in golden["rehash"]], dtype=np.uint64)
    assert [str(int(h)) for h in oracle_hash(col, m, seeds)] == [h for _, h in golden["rehash"]]


def test_contains_matches_reference_vectors_and_bytes_find(golden):
    m = golden["max_len"]
    col = varchar.pack(golden["values"], m)
    for needle_hex, want in golden["contains"]:
        needle = bytes.fromhex(needle_hex)
        got = oracle_contains(col, m, needle)
        assert got.tolist() == want, needle
        assert got.tolist() == [int(needle in v) for v in golden["values"]]


def strncmp_equal(a: bytes, b: bytes) -> bool:
    """varchar::operator==: equal lengths, then strncmp over that length,
    which ends at the first NUL the two share."""
    if len(a) != len(b):
        return False
    cut = a.find(b"\x00")
    return a == b if cut < 0 else (a[:cut + 1] == b[:cut + 1])


def test_equality_matches_reference_vectors_and_strncmp_rule(golden):
    m = golden["max_len"]
    vals = golden["values"]
    n = len(vals)
    col = varchar.pack(vals, m)
    for j in range(n):
        other = varchar.pack([vals[j]] * n, m)
        got = oracle_eq(col, m, other, m)
        assert got.tolist() == [golden["eq"][i * n + j] for i in range(n)], j
        assert got.tolist() == [int(strncmp_equal(vals[i], vals[j])) for i in range(n)]
    # bytes after a NUL both strings share are not compared
    assert strncmp_equal(b"ab\x00cd", b"ab\x00ce") and not strncmp_equal(b"abcd", b"abce")


@pytest.mark.skipif(not ref.available(), reason="reference sources not present")
def test_oracle_against_live_reference_varchar():
    L = ref.lib()
    rng = np.random.default_rng(5)
    for m in (1, 7, 8, 9, 55, 255):
        vals = [bytes(rng.integers(0, 4, int(rng.integers(0, m + 1)), dtype=np.uint8) + 97) for _ in range(300)]
        col = varchar.pack(vals, m)
        h = oracle_hash(col, m)
        for i in range(0, 300, 7):
            assert int(h[i]) == L.ref_hash_varchar(col.ctypes.data, i, m)
        for needle in (b"", b"ab", b"dcba", b"a" * 5):
            got = oracle_contains(col, m, needle)
            for i in range(0, 300, 11):
                assert got[i] == L.ref_varchar_contains(col.ctypes.data, i, m, needle, len(needle))


def test_slot_layout_and_unpack():
    col = varchar.pack(["", "abc", b"\xff\x00z"], 5)
    assert col.shape == (3, 6) and col.dtype == np.uint8
    assert col[1].tolist() == [3, 97, 98, 99, 0, 0]
    assert varchar.unpack(col) == [b"", b"abc", b"\xff\x00z"]
    with pytest.raises(ValueError):
        varchar.pack(["toolong"], 3)
    with pytest.raises(ValueError):
        varchar.pack(["x"], 256)


def test_importer_char_and_varchar_casts():
    assert tbl.cast_char("  BUILDING", 10) == b"BUILDING"       # Char<n>: leading spaces dropped
    assert tbl.cast_char("AUTO MOBILE ", 12) == b"AUTO MOBILE "   # trailing ones kept
    assert tbl.cast_varchar("  green  ", 10) == b"  green  "      # Varchar<n>: as it is
    with pytest.raises(tbl.FormatError):
        tbl.cast_char("ABCDEFGHIJK", 10)
    with pytest.raises(tbl.FormatError):
        tbl.cast_varchar("x" * 56, 55)


def test_contains_predicate_feeds_a_join_cpu(tmp_path):
    """p_name contains 'green' (the oracle's vec_contains) as a u8 build-side
    predicate of a lineitem-part join: the oracle's answer equals the one
    computed straight from the text."""
    from tbl_data import green_revenue, green_revenue_from_text, write_lineitem, write_part
    write_lineitem(tmp_path / "lineitem.tbl", 3000, seed=21)
    write_part(tmp_path / "part.tbl", 1000, seed=22)
    pa, _ = tbl.load_table(str(tmp_path), "part", ["p_partkey", "p_name"])
    li, lb = tbl.load_table(str(tmp_path), "lineitem",
                            ["l_partkey", "l_extendedprice", "l_discount", "l_returnflag"])
    green = oracle_contains(pa["p_name"], 55, b"green")
    assert 0 < green.sum() < len(green)
    got = oracle.sorted_rows(oracle.bruteforce(green_revenue(li, pa["p_partkey"], green, lb).desc()))
    assert got == green_revenue_from_text(tmp_path / "lineitem.tbl", tmp_path / "part.tbl")
