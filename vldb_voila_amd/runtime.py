"""ctypes binding of libvoila_hip.so.  Loading fails loudly when the library
is missing: there is no CPU fallback on the product path."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libvoila_hip.so")
_LIB = None


class VoilaHipError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{abi.STATUS_NAMES.get(code, code)}: {msg}")
        self.code = code


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        vp, sz, i32, i64 = C.c_void_p, C.c_size_t, C.c_int32, C.c_int64
        L.vh_last_error.restype = C.c_char_p
        L.vh_init.argtypes = [C.c_int]
        L.vh_malloc.argtypes = [C.POINTER(vp), sz]
        L.vh_free.argtypes = [vp]
        L.vh_memcpy_h2d.argtypes = [vp, vp, sz]
        L.vh_memcpy_d2h.argtypes = [vp, vp, sz]
        L.vh_stream_create.argtypes = [C.POINTER(vp)]
        L.vh_stream_destroy.argtypes = [vp]
        L.vh_stream_synchronize.argtypes = [vp]
        L.vh_event_create.argtypes = [C.POINTER(vp)]
        L.vh_event_destroy.argtypes = [vp]
        L.vh_event_record.argtypes = [vp, vp]
        L.vh_event_elapsed_ms.argtypes = [C.POINTER(C.c_float), vp, vp]
        L.vh_query_new.argtypes = [C.POINTER(vp), vp, vp]
        L.vh_query_run.argtypes = [vp, vp]
        L.vh_query_run_phase.argtypes = [vp, vp, C.c_int]
        L.vh_query_result.argtypes = [vp, vp, C.c_char_p, sz, C.POINTER(sz), C.POINTER(i64)]
        L.vh_query_reset.argtypes = [vp]
        L.vh_query_free.argtypes = [vp]
        L.vh_query_info.argtypes = [vp, C.POINTER(i32), C.POINTER(i32), C.POINTER(i64)]
        L.vh_query_kernel.argtypes = [vp, C.POINTER(i32), C.c_char_p, C.c_size_t, C.POINTER(C.c_size_t)]
        L.vh_query_join.argtypes = [vp, i32, C.POINTER(i32), C.POINTER(i64)]
        L.vh_type_from_minmax.argtypes = [C.c_double, C.c_double]
        _LIB = L
    return _LIB


def check(rc):
    if rc != 0:
        raise VoilaHipError(rc, lib().vh_last_error().decode())
    return rc


_INIT = False


def init(device=0):
    global _INIT
    if not _INIT:
        check(lib().vh_init(device))
        _INIT = True


class DeviceBuffer:
    def __init__(self, nbytes):
        init()
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        check(lib().vh_malloc(C.byref(p), max(self.nbytes, 1)))
        self.ptr = p.value

    @classmethod
    def from_numpy(cls, arr: np.ndarray):
        arr = np.ascontiguousarray(arr)
        b = cls(arr.nbytes)
        if arr.nbytes:
            check(lib().vh_memcpy_h2d(b.ptr, arr.ctypes.data, arr.nbytes))
        return b

    def copy_from(self, arr: np.ndarray):
        """Overwrite the buffer's first arr.nbytes bytes with `arr`."""
        arr = np.ascontiguousarray(arr)
        if arr.nbytes > self.nbytes:
            raise ValueError("array larger than the buffer")
        if arr.nbytes:
            check(lib().vh_memcpy_h2d(self.ptr, arr.ctypes.data, arr.nbytes))

    def to_numpy(self, dtype, count):
        out = np.empty(count, dtype=dtype)
        if out.nbytes:
            check(lib().vh_memcpy_d2h(out.ctypes.data, self.ptr, out.nbytes))
        return out

    def free(self):
        if self.ptr:
            lib().vh_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def upload(arrays: dict) -> dict:
    """Device copies of a dict of numpy columns; returns name -> DeviceBuffer."""
    return {k: DeviceBuffer.from_numpy(v) for k, v in arrays.items()}


def addrs(bufs: dict) -> dict:
    return {k: v.ptr for k, v in bufs.items()}


class Stream:
    def __init__(self):
        init()
        p = C.c_void_p()
        check(lib().vh_stream_create(C.byref(p)))
        self.ptr = p.value

    def synchronize(self):
        check(lib().vh_stream_synchronize(self.ptr))

    def __del__(self):
        try:
            if self.ptr:
                lib().vh_stream_destroy(self.ptr)
        except Exception:
            pass


class Event:
    def __init__(self):
        init()
        p = C.c_void_p()
        check(lib().vh_event_create(C.byref(p)))
        self.ptr = p.value

    def record(self, stream=None):
        check(lib().vh_event_record(self.ptr, stream.ptr if stream else None))

    def elapsed_ms(self, end: "Event") -> float:
        ms = C.c_float()
        check(lib().vh_event_elapsed_ms(C.byref(ms), self.ptr, end.ptr))
        return ms.value

    def __del__(self):
        try:
            if self.ptr:
                lib().vh_event_destroy(self.ptr)
        except Exception:
            pass


def make_config(group_mode=-1, grid=0, table_shift=0, tile_shift=0, join_layout=0, kernel=0, unroll=0, merge=0,
                part_shift=0, block=0, radix=0):
    c = abi.vh_config()
    c.part_shift = part_shift
    c.block = block
    c.radix = radix
    c.group_mode = group_mode
    c.grid = grid
    c.table_shift = table_shift
    c.tile_shift = tile_shift
    c.join_layout = join_layout
    c.kernel = kernel
    c.unroll = unroll
    c.merge = merge
    return c


class Query:
    """voila_query_new / run / reset / free over a vh_pipeline_desc whose
    column pointers are device addresses."""

    def __init__(self, desc: abi.vh_pipeline_desc, config: abi.vh_config | None = None):
        init()
        self.desc = desc
        self.config = config
        p = C.c_void_p()
        check(lib().vh_query_new(C.byref(p), C.byref(desc),
                                 C.byref(config) if config is not None else None))
        self.ptr = p.value

    def run(self, stream: Stream | None = None):
        check(lib().vh_query_run(self.ptr, stream.ptr if stream else None))

    def run_phase(self, phase: int, stream: Stream | None = None):
        check(lib().vh_query_run_phase(self.ptr, stream.ptr if stream else None, phase))

    def result(self, stream: Stream | None = None) -> str:
        n = C.c_size_t(0)
        rows = C.c_int64(0)
        s = stream.ptr if stream else None
        check(lib().vh_query_result(self.ptr, s, None, 0, C.byref(n), C.byref(rows)))
        buf = C.create_string_buffer(n.value + 1)
        check(lib().vh_query_result(self.ptr, s, buf, n.value + 1, C.byref(n), C.byref(rows)))
        return buf.value.decode()

    def info(self):
        m, g, n = C.c_int32(), C.c_int32(), C.c_int64()
        check(lib().vh_query_info(self.ptr, C.byref(m), C.byref(g), C.byref(n)))
        return {"group_mode": m.value, "grid": g.value, "groups": n.value}

    def join(self, j=0):
        """(entry bytes, entries) of join j's table: 1/2/4 direct, 8 hash."""
        w, n = C.c_int32(), C.c_int64()
        check(lib().vh_query_join(self.ptr, j, C.byref(w), C.byref(n)))
        return w.value, n.value

    def kernel(self):
        """(enum vh_kernel in use, generated HIP source or '')."""
        k, n = C.c_int32(), C.c_size_t()
        check(lib().vh_query_kernel(self.ptr, C.byref(k), None, 0, C.byref(n)))
        buf = C.create_string_buffer(n.value + 1)
        check(lib().vh_query_kernel(self.ptr, None, buf, n.value + 1, C.byref(n)))
        return k.value, buf.value.decode()

    def probe_events(self, start: "Event | None", stop: "Event | None"):
        """Events later runs stamp with the generated probe kernel's own
        start and end (None, None: stop)."""
        L = lib()
        L.vh_query_probe_events.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        check(L.vh_query_probe_events(self.ptr, start.ptr if start else None, stop.ptr if stop else None))

    def reset(self):
        check(lib().vh_query_reset(self.ptr))

    def free(self):
        if self.ptr:
            lib().vh_query_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def run_pipeline(pipeline_builder, config=None) -> str:
    """Convenience: build desc from a Pipeline whose columns are device
    addresses, run once, return the result text."""
    q = Query(pipeline_builder.desc(), config)
    try:
        q.run()
        return q.result()
    finally:
        q.free()


class Batch:
    """vh_batch: many pipelines in one persistent launch (optionally replayed
    from a captured hipGraph)."""

    def __init__(self, descs, configs=None, grid=0):
        init()
        L = lib()
        L.vh_batch_new.argtypes = [C.POINTER(C.c_void_p), C.c_int32, C.c_void_p, C.c_void_p, C.c_int32]
        L.vh_batch_run.argtypes = [C.c_void_p, C.c_void_p]
        L.vh_batch_run_graph.argtypes = [C.c_void_p, C.c_void_p]
        L.vh_batch_result.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_char_p, C.c_size_t,
                                      C.POINTER(C.c_size_t), C.POINTER(C.c_int64)]
        L.vh_batch_info.argtypes = [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                    C.POINTER(C.c_int64)]
        L.vh_batch_free.argtypes = [C.c_void_p]
        n = len(descs)
        arr = (abi.vh_pipeline_desc * n)(*descs)
        cfg = None
        if configs is not None:
            cfg = (abi.vh_config * n)(*configs)
        self._keep = (arr, cfg)
        p = C.c_void_p()
        check(L.vh_batch_new(C.byref(p), n, C.addressof(arr), C.addressof(cfg) if cfg is not None else None,
                             grid))
        self.ptr = p.value
        self.n = n

    def run(self, stream: Stream | None = None):
        check(lib().vh_batch_run(self.ptr, stream.ptr if stream else None))

    def run_graph(self, stream: Stream):
        check(lib().vh_batch_run_graph(self.ptr, stream.ptr))

    def result(self, i, stream: Stream | None = None) -> str:
        n = C.c_size_t(0)
        rows = C.c_int64(0)
        s = stream.ptr if stream else None
        check(lib().vh_batch_result(self.ptr, i, s, None, 0, C.byref(n), C.byref(rows)))
        buf = C.create_string_buffer(n.value + 1)
        check(lib().vh_batch_result(self.ptr, i, s, buf, n.value + 1, C.byref(n), C.byref(rows)))
        return buf.value.decode()

    def info(self):
        it, g, l = C.c_int32(), C.c_int32(), C.c_int64()
        check(lib().vh_batch_info(self.ptr, C.byref(it), C.byref(g), C.byref(l)))
        return {"items": it.value, "grid": g.value, "lds_bytes": l.value}

    def free(self):
        if self.ptr:
            lib().vh_batch_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
