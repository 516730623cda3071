"""vh_vec_seltrue_noalloc, the selection entry point that allocates nothing
(its tile counts live in a static device array), against vh_vec_seltrue
through the C ABI: two calls in a row, the second on a smaller input, dense
and with a selection vector, and its size limit."""
import ctypes as C

import numpy as np
import pytest

from vldb_voila_amd import runtime
from vldb_voila_amd.runtime import DeviceBuffer

pytestmark = pytest.mark.gpu

MAX_POSITIONS = 131071 * 4096


def _bind():
    L = runtime.lib()
    vp, i64 = C.c_void_p, C.c_int64
    for name in ("vh_vec_seltrue", "vh_vec_seltrue_noalloc"):
        getattr(L, name).argtypes = [vp, i64, vp, vp, vp, vp]
        getattr(L, name).restype = C.c_int
    return L


def _select(fn, sel, n, pred, stream):
    res = DeviceBuffer(max(n, 1) * 4)
    cnt = DeviceBuffer(16)
    runtime.check(fn(sel.ptr if sel is not None else None, n, res.ptr, pred.ptr, cnt.ptr, stream.ptr))
    stream.synchronize()
    k = int(cnt.to_numpy(np.int32, 1)[0])
    return res.to_numpy(np.int32, k)


@pytest.mark.parametrize("with_sel", [False, True])
def test_noalloc_matches_seltrue_on_two_calls(with_sel):
    print(f"seltrue_noalloc with_sel={with_sel}", flush=True)
    runtime.init(0)
    L = _bind()
    stream = runtime.Stream()
    rng = np.random.default_rng(11)
    for n, density in ((5_000_011, 0.4), (70_001, 0.8)):
        dom = 2 * n if with_sel else n
        pred = DeviceBuffer.from_numpy((rng.random(dom) < density).astype(np.uint8))
        sel = None
        if with_sel:
            sel = DeviceBuffer.from_numpy(np.sort(rng.choice(dom, size=n, replace=False)).astype(np.int32))
        want = _select(L.vh_vec_seltrue, sel, n, pred, stream)
        got = _select(L.vh_vec_seltrue_noalloc, sel, n, pred, stream)
        assert len(want) > 0
        assert np.array_equal(got, want)


def test_noalloc_empty_input_and_limit():
    print("seltrue_noalloc empty / limit", flush=True)
    runtime.init(0)
    L = _bind()
    stream = runtime.Stream()
    cnt = DeviceBuffer(16)
    cnt.copy_from(np.array([7, 0, 0, 0], dtype=np.int32))
    runtime.check(L.vh_vec_seltrue_noalloc(None, 0, None, None, cnt.ptr, stream.ptr))
    stream.synchronize()
    assert int(cnt.to_numpy(np.int32, 1)[0]) == 0
    # one position past the static scratch: refused before anything runs
    dummy = DeviceBuffer(16)
    rc = L.vh_vec_seltrue_noalloc(None, MAX_POSITIONS + 1, dummy.ptr, dummy.ptr, cnt.ptr, stream.ptr)
    assert rc == 5   # VH_ERR_CAPACITY
    assert b"static scratch" in L.vh_last_error()
