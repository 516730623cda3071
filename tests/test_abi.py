"""The C ABI library loads and exports every entry point include/voila_hip.h
declares; descriptor layouts agree between C and ctypes.  No GPU needed."""
import ctypes as C
import os
import re

from vldb_voila_amd import abi, runtime

HEADER = os.path.join(os.path.dirname(os.path.dirname(__file__)), "include", "voila_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(vh_[a-z_0-9]+)\s*\(", text, flags=re.M)))


def test_header_declares_entry_points():
    names = declared_functions()
    assert "vh_query_new" in names and "vh_query_run" in names and len(names) > 15


def test_library_exports_every_declared_symbol():
    L = runtime.lib()
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing


def test_descriptor_sizes():
    # fixed-layout C structs: 8-byte aligned, no hidden padding surprises
    assert C.sizeof(abi.vh_column) == 32
    assert C.sizeof(abi.vh_pred) == 16
    assert C.sizeof(abi.vh_expr) == 24
    assert C.sizeof(abi.vh_table_desc) == 16 + abi.VH_MAX_COLS * 32 + abi.VH_MAX_PREDS * 16
    assert C.sizeof(abi.vh_config) == 48


def test_type_from_minmax_matches_oracle():
    import oracle
    L = runtime.lib()
    for lo, hi in [(0, 255), (-1, 127), (0, 2**32 - 1), (-5, 2**40), (0, 2**63), (-1, 2**64), (0, 1e21)]:
        assert L.vh_type_from_minmax(lo, hi) == oracle.lib().vo_type_from_minmax(lo, hi)


def test_errors_without_gpu_are_reported():
    L = runtime.lib()
    # no device in the build container: vh_init fails loudly, never silently
    rc = L.vh_init(0)
    if rc != 0:
        assert L.vh_last_error()
