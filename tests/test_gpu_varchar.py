"""Varchar statements on the GPU (vh_vec_hash_varchar / rehash / contains /
cmp_varchar through the C ABI) against the oracle, bit-exact: dense ranges
(LDS-staged tiles), selection vectors, column bases off dword alignment,
every max_len class, empty inputs; and p_name contains 'green' as the
build-side predicate of a lineitem-part join over imported .tbl columns."""
import numpy as np
import pytest

import oracle
from vldb_voila_amd import abi, runtime, tbl, varchar
from vldb_voila_amd.runtime import DeviceBuffer

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def device():
    print("gpu varchar: init device 0", flush=True)
    runtime.init(0)


def strings(rng, n, max_len):
    """Short alphabets so that contains and equality hit often; lengths 0..max_len."""
    lens = rng.integers(0, max_len + 1, n)
    lens[: min(n, 4)] = [0, max_len, min(8, max_len), 1][: min(n, 4)]
    return [bytes(rng.integers(97, 100, int(k), dtype=np.uint8)) for k in lens]


def upload_at(col: np.ndarray, shift: int):
    """Device copy of a slot column starting `shift` bytes into its buffer."""
    flat = col.reshape(-1)
    buf = DeviceBuffer(flat.nbytes + 8)
    if shift:
        padded = np.concatenate([np.full(shift, 0xEE, dtype=np.uint8), flat])
        buf.copy_from(padded)
    else:
        buf.copy_from(flat)
    return buf, buf.ptr + shift


def ora(fn, n, dtype, *args):
    res = np.zeros(n, dtype=dtype)
    fn(None, n, res.ctypes.data, *args)
    return res


@pytest.mark.parametrize("max_len", [1, 7, 8, 25, 55, 199, 255])
def test_varchar_statements_dense_and_unaligned(max_len):
    print(f"varchar statements dense max_len={max_len}", flush=True)
    rng = np.random.default_rng(max_len)
    n = 3001
    vals = strings(rng, n, max_len)
    col = varchar.pack(vals, max_len)
    L = oracle.lib()
    want_h = ora(lambda s, k, r: L.vo_vec_hash_varchar(s, k, r, None, col.ctypes.data, max_len), n, np.uint64)
    prev = rng.integers(0, 2**63, n, dtype=np.int64).astype(np.uint64)
    want_rh = ora(lambda s, k, r: L.vo_vec_hash_varchar(s, k, r, prev.ctypes.data, col.ctypes.data, max_len),
                  n, np.uint64)
    dprev = DeviceBuffer.from_numpy(prev)
    for shift in (0, 1, 3):
        buf, ptr = upload_at(col, shift)
        assert np.array_equal(varchar.hash(ptr, n, max_len).to_numpy(np.uint64, n), want_h), shift
        assert np.array_equal(varchar.hash(ptr, n, max_len, prev=dprev).to_numpy(np.uint64, n), want_rh), shift
        for needle in (b"", b"a", b"ab", b"cab", vals[1][: max(1, max_len // 2)], b"d", b"a" * min(max_len + 1, 255)):
            want = ora(lambda s, k, r: L.vo_vec_contains(s, k, r, col.ctypes.data, max_len, needle, len(needle)),
                       n, np.uint8)
            got = varchar.contains(ptr, n, max_len, needle).to_numpy(np.uint8, n)
            assert np.array_equal(got, want), (shift, needle)
        for cst in (b"", vals[2], vals[3], b"abc"[:max_len]):
            for cmp in ("==", "!="):
                want = ora(lambda s, k, r: L.vo_vec_cmp_varchar(abi.CMP_NAMES[cmp], s, k, r, col.ctypes.data, max_len,
                                                               None, 0, cst, len(cst)), n, np.uint8)
                got = varchar.compare(cmp, ptr, n, max_len, cst).to_numpy(np.uint8, n)
                assert np.array_equal(got, want), (shift, cst, cmp)


def test_varchar_statements_with_selection_and_second_column():
    print("varchar statements with sel and column operands", flush=True)
    rng = np.random.default_rng(11)
    n, m1, m2 = 20_000, 12, 9
    a = strings(rng, n, m1)
    b = [v[:m2] if rng.random() < 0.5 else bytes(rng.integers(97, 100, int(rng.integers(0, m2 + 1)), dtype=np.uint8))
         for v in a]
    c1, c2 = varchar.pack(a, m1), varchar.pack(b, m2)
    d1, d2 = DeviceBuffer.from_numpy(c1), DeviceBuffer.from_numpy(c2)
    pos = np.sort(rng.choice(n, n // 3, replace=False)).astype(np.int32)
    sel = varchar.Selection(pos)
    L = oracle.lib()
    # untouched positions keep what the buffer held
    out = DeviceBuffer.from_numpy(np.full(n, 7, dtype=np.uint8))
    varchar.compare("==", d1.ptr, n, m1, d2.ptr, m2, sel=sel, out=out)
    want = np.full(n, 7, dtype=np.uint8)
    L.vo_vec_cmp_varchar(abi.VH_EQ, pos.ctypes.data, len(pos), want.ctypes.data, c1.ctypes.data, m1,
                         c2.ctypes.data, m2, None, 0)
    got = out.to_numpy(np.uint8, n)
    assert np.array_equal(got, want)
    assert 0 < int((got[pos] == 1).sum()) < len(pos)
    hv = DeviceBuffer.from_numpy(np.zeros(n, dtype=np.uint64))
    varchar.hash(d1.ptr, n, m1, sel=sel, out=hv)
    wh = np.zeros(n, dtype=np.uint64)
    L.vo_vec_hash_varchar(pos.ctypes.data, len(pos), wh.ctypes.data, None, c1.ctypes.data, m1)
    assert np.array_equal(hv.to_numpy(np.uint64, n), wh)
    cv = DeviceBuffer.from_numpy(np.zeros(n, dtype=np.uint8))
    varchar.contains(d1.ptr, n, m1, b"bc", sel=sel, out=cv)
    wc = np.zeros(n, dtype=np.uint8)
    L.vo_vec_contains(pos.ctypes.data, len(pos), wc.ctypes.data, c1.ctypes.data, m1, b"bc", 2)
    assert np.array_equal(cv.to_numpy(np.uint8, n), wc)
    # empty inputs and bad arguments
    varchar.contains(d1.ptr, 0, m1, b"a")
    varchar.hash(d1.ptr, 0, m1, sel=varchar.Selection(np.zeros(0, dtype=np.int32)))
    with pytest.raises(runtime.VoilaHipError):
        varchar.contains(d1.ptr, n, 0, b"a")
    with pytest.raises(runtime.VoilaHipError):
        varchar.contains(d1.ptr, n, m1, b"x" * 256)
    with pytest.raises(ValueError):
        varchar.compare("<", d1.ptr, n, m1, b"a")


def test_contains_predicate_feeds_a_join(tmp_path):
    """p_name contains 'green' computed on the GPU, used as the u8
    build-side predicate of a lineitem-part join run by the fused kernel."""
    print("varchar: green parts join on the GPU", flush=True)
    from tbl_data import green_revenue, green_revenue_from_text, write_lineitem, write_part
    write_lineitem(tmp_path / "lineitem.tbl", 20_000, seed=31)
    write_part(tmp_path / "part.tbl", 1000, seed=32)
    pa, _ = tbl.load_table(str(tmp_path), "part", ["p_partkey", "p_name"])
    li, lb = tbl.load_table(str(tmp_path), "lineitem",
                            ["l_partkey", "l_extendedprice", "l_discount", "l_returnflag"])
    dname = DeviceBuffer.from_numpy(pa["p_name"])
    green = varchar.contains(dname.ptr, len(pa["p_name"]), 55, "green")
    host_green = green.to_numpy(np.uint8, len(pa["p_name"]))
    assert 0 < host_green.sum() < len(host_green)
    dkeys = DeviceBuffer.from_numpy(pa["p_partkey"])
    dli = runtime.upload(li)
    p = green_revenue(li, pa["p_partkey"], host_green, lb,
                      addrs={"p_partkey": dkeys.ptr, "lineitem": runtime.addrs(dli)}, green_addr=green.ptr)
    want = green_revenue_from_text(tmp_path / "lineitem.tbl", tmp_path / "part.tbl")
    for kernel in (abi.VH_KERNEL_GENERIC, abi.VH_KERNEL_JIT):
        got = oracle.sorted_rows(runtime.run_pipeline(p, runtime.make_config(kernel=kernel)))
        assert got == want, kernel
