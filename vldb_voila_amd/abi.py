"""ctypes mirror of include/voila_hip.h (descriptor structs and enums)."""
from __future__ import annotations

import ctypes as C

import numpy as np

VH_INVALID, VH_U8, VH_I8, VH_U16, VH_I16, VH_U32, VH_I32, VH_U64, VH_I64, VH_I128 = range(10)
VH_LT, VH_LE, VH_GT, VH_GE, VH_EQ, VH_NE = range(6)
VH_E_COL, VH_E_CONST, VH_E_ADD, VH_E_SUB, VH_E_MUL = range(5)
VH_A_SUM, VH_A_COUNT = range(2)
VH_JOIN01, VH_JOINN = range(2)

VH_MAX_COLS = 8
VH_MAX_PREDS = 8
VH_MAX_JOINS = 4
VH_MAX_PAYLOAD = 4
VH_MAX_JOIN_KEYS = 3
VH_MAX_KEYS = 4
VH_MAX_AGGRS = 8
VH_MAX_EXPR = 32

STATUS_NAMES = {
    0: "VH_OK", 1: "VH_ERR_INVALID", 2: "VH_ERR_UNSUPPORTED", 3: "VH_ERR_NOMEM",
    4: "VH_ERR_HIP", 5: "VH_ERR_CAPACITY", 6: "VH_ERR_STATE", 7: "VH_ERR_NODEVICE",
}

CMP_NAMES = {"<": VH_LT, "<=": VH_LE, ">": VH_GT, ">=": VH_GE, "==": VH_EQ, "!=": VH_NE,
             "lt": VH_LT, "le": VH_LE, "gt": VH_GT, "ge": VH_GE, "eq": VH_EQ, "ne": VH_NE}

_DTYPE_TO_VH = {
    np.dtype(np.uint8): VH_U8, np.dtype(np.int8): VH_I8,
    np.dtype(np.uint16): VH_U16, np.dtype(np.int16): VH_I16,
    np.dtype(np.uint32): VH_U32, np.dtype(np.int32): VH_I32,
    np.dtype(np.uint64): VH_U64, np.dtype(np.int64): VH_I64,
}
VH_TO_DTYPE = {v: k for k, v in _DTYPE_TO_VH.items()}
VH_WIDTH = {VH_U8: 1, VH_I8: 1, VH_U16: 2, VH_I16: 2, VH_U32: 4, VH_I32: 4,
            VH_U64: 8, VH_I64: 8, VH_I128: 16}


def vh_type_of(dtype) -> int:
    return _DTYPE_TO_VH[np.dtype(dtype)]


class vh_column(C.Structure):
    _fields_ = [("type", C.c_int32), ("reserved", C.c_int32), ("data", C.c_void_p),
                ("min", C.c_int64), ("max", C.c_int64)]


class vh_pred(C.Structure):
    _fields_ = [("cmp", C.c_int32), ("col", C.c_int32), ("value", C.c_int64)]


class vh_expr(C.Structure):
    _fields_ = [("op", C.c_int32), ("a", C.c_int32), ("b", C.c_int32),
                ("reserved", C.c_int32), ("value", C.c_int64)]


class vh_table_desc(C.Structure):
    _fields_ = [("rows", C.c_int64), ("ncols", C.c_int32), ("npreds", C.c_int32),
                ("cols", vh_column * VH_MAX_COLS), ("preds", vh_pred * VH_MAX_PREDS)]


class vh_join_desc(C.Structure):
    _fields_ = [("build", vh_table_desc), ("build_key", C.c_int32), ("probe_key", C.c_int32),
                ("variant", C.c_int32), ("npayload", C.c_int32),
                ("payload", C.c_int32 * VH_MAX_PAYLOAD), ("nkeys", C.c_int32),
                ("build_keys", C.c_int32 * (VH_MAX_JOIN_KEYS - 1)),
                ("probe_keys", C.c_int32 * (VH_MAX_JOIN_KEYS - 1))]


class vh_aggr_desc(C.Structure):
    _fields_ = [("fn", C.c_int32), ("expr", C.c_int32)]


class vh_pipeline_desc(C.Structure):
    _fields_ = [("probe", vh_table_desc), ("njoins", C.c_int32), ("nkeys", C.c_int32),
                ("naggrs", C.c_int32), ("nexpr", C.c_int32),
                ("joins", vh_join_desc * VH_MAX_JOINS), ("keys", C.c_int32 * VH_MAX_KEYS),
                ("aggrs", vh_aggr_desc * VH_MAX_AGGRS), ("expr", vh_expr * VH_MAX_EXPR)]


VH_GROUP_AUTO, VH_GROUP_REG, VH_GROUP_LDS, VH_GROUP_GLOBAL, VH_GROUP_HASH, VH_GROUP_LDSHASH = -1, 0, 1, 2, 3, 4
VH_GROUP_DENSE = 5
VH_GROUP_PART = 6

# enum vh_join_layout
VH_LAYOUT_AUTO = 0
VH_LAYOUT_HASH = 1
VH_LAYOUT_DIRECT = 2

# enum vh_kernel
VH_KERNEL_AUTO = 0
VH_KERNEL_GENERIC = 1
VH_KERNEL_JIT = 2
VH_KERNEL_JIT_LAZY = 3
VH_KERNEL_JIT_PIPE = 4


class vh_config(C.Structure):
    _fields_ = [("group_mode", C.c_int32), ("grid", C.c_int32), ("table_shift", C.c_int32),
                ("tile_shift", C.c_int32), ("join_layout", C.c_int32), ("kernel", C.c_int32),
                ("unroll", C.c_int32), ("merge", C.c_int32), ("part_shift", C.c_int32),
                ("block", C.c_int32), ("radix", C.c_int32), ("reserved", C.c_int32)]
