"""Run log in the reference's format (CPU): the table has the reference's
columns, blends split as voila_tools.parse_blend splits them, and the
base-flavor selection orders blends by average time and skips failed runs."""
import sqlite3

from vldb_voila_amd import runlog, runtime

# column names of compiler.cpp:105-111, in order
REFERENCE_COLUMNS = ["id", "t", "mode", "query", "result", "rep", "time_ms", "tuples", "cycles", "instructions",
                     "L1misses", "LLCmisses", "branchmisses", "taskclock", "threads", "cgen_ms", "ccomp_ms",
                     "default_blend", "key_check_blend", "aggregates_blend", "backend", "scale_factor",
                     "pipeline_flavor", "full_blend"]


def parse_blend(blend):
    """The split voila_tools.parse_blend does (voila_tools.py:2-8)."""
    return dict(pair.split("=") for pair in blend.split(","))


def test_table_has_the_reference_columns(tmp_path):
    conn = runlog.open_log(str(tmp_path / "runs.db"))
    cols = [r[1] for r in conn.execute("PRAGMA table_info(runs)")]
    assert cols == REFERENCE_COLUMNS


def test_blend_string_round_trips_through_the_reference_split():
    c = runtime.make_config(group_mode=4, grid=512, join_layout=2, kernel=2, unroll=8)
    b = parse_blend(runlog.blend_string(c))
    assert b == {"group_mode": "lds_hash", "join_layout": "direct", "table_shift": "0", "grid": "512",
                 "kernel": "jit", "unroll": "8"}


def test_base_flavors_orders_by_average_and_skips_failures(tmp_path):
    conn = runlog.open_log(str(tmp_path / "runs.db"))
    fast, slow = "group_mode=lds,grid=512", "group_mode=hash,grid=256"
    for rep, t in enumerate([2.0, 4.0]):
        runlog.record(conn, query="q", blend=slow, time_ms=t, tuples=10, rep=rep)
    for rep, t in enumerate([1.0, 3.0]):
        runlog.record(conn, query="q", blend=fast, time_ms=t, tuples=10, rep=rep)
    runlog.record(conn, query="q", blend="group_mode=part", time_ms=None, tuples=10, result="CRASH")
    runlog.record(conn, query="other", blend=fast, time_ms=0.1, tuples=10)
    rows = runlog.base_flavors(conn, "q")
    assert [r[3] for r in rows] == [fast, slow]
    assert rows[0][:3] == (1.0, 2.0, 3.0) and rows[1][:3] == (2.0, 3.0, 4.0)
    assert isinstance(conn, sqlite3.Connection)
