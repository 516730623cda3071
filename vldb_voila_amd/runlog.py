"""The explorer's run log in the reference's format.

The reference records every timed run in an SQLite table `runs`
(compiler.cpp:100-122, record_run at :124-255), and its analyses read it
back, e.g. the fastest flavors of a query (print_base_flavors.py:24-45: min /
avg / max of time_ms per default_blend).  This module writes the GPU
explorer's timings into a table with the same columns, so those analyses run
unchanged on them: `default_blend` holds the variant as `key=value` pairs
separated by commas, the format voila_tools.parse_blend splits, and
`backend` is "hip".
"""
from __future__ import annotations

import sqlite3

from . import abi

# The reference's table, column for column (compiler.cpp:105-111).
RUNS_DDL = (
    "CREATE TABLE IF NOT EXISTS runs(id INTEGER PRIMARY KEY AUTOINCREMENT, "
    "t TIMESTAMP, mode TEXT, query TEXT, result TEXT, rep INTEGER, time_ms REAL NULL, "
    "tuples REAL NULL, cycles REAL NULL, instructions REAL NULL, L1misses REAL NULL, "
    "LLCmisses REAL NULL, branchmisses REAL NULL, taskclock REAL NULL, threads INTEGER, "
    "cgen_ms REAL NULL, ccomp_ms REAL NULL, default_blend TEXT, key_check_blend TEXT, "
    "aggregates_blend TEXT, backend TEXT, scale_factor INTEGER, pipeline_flavor TEXT, "
    "full_blend TEXT)")

_GROUP_NAMES = {-1: "auto", 0: "reg", 1: "lds", 2: "global", 3: "hash", 4: "lds_hash", 5: "dense", 6: "part"}
_LAYOUT_NAMES = {0: "auto", 1: "hash", 2: "direct"}
_KERNEL_NAMES = {0: "auto", 1: "generic", 2: "jit", 3: "jit_lazy", 4: "jit"}


def open_log(path: str) -> sqlite3.Connection:
    conn = sqlite3.connect(path)
    conn.execute(RUNS_DDL)
    return conn


def blend_string(cfg: abi.vh_config) -> str:
    """A variant as the reference writes a blend: `key=value,...`."""
    return ",".join([
        f"group_mode={_GROUP_NAMES.get(cfg.group_mode, str(cfg.group_mode))}",
        f"join_layout={_LAYOUT_NAMES.get(cfg.join_layout, str(cfg.join_layout))}",
        f"table_shift={cfg.table_shift}",
        f"grid={cfg.grid}",
        f"kernel={_KERNEL_NAMES.get(cfg.kernel, str(cfg.kernel))}",
        f"unroll={cfg.unroll}",
    ])


def record(conn: sqlite3.Connection, *, query: str, blend: str, time_ms: float | None, tuples: float | None,
           rep: int = 0, mode: str = "explore", result: str = "OK", threads: int = 1, scale_factor: int = 0,
           pipeline_flavor: str = "", full_blend: str = "") -> None:
    """One run.  A variant that does not apply to the pipeline is recorded
    with result 'CRASH' and no time, as the reference records a run that
    failed; its analyses skip such rows."""
    conn.execute(
        "INSERT INTO runs(t, mode, query, result, rep, time_ms, tuples, threads, default_blend, "
        "key_check_blend, aggregates_blend, backend, scale_factor, pipeline_flavor, full_blend) "
        "VALUES (datetime('now','localtime'), ?, ?, ?, ?, ?, ?, ?, ?, '', '', 'hip', ?, ?, ?)",
        (mode, query, result, rep, time_ms, tuples, threads, blend, scale_factor, pipeline_flavor, full_blend))


def base_flavors(conn: sqlite3.Connection, query: str, mode: str = "explore", threads: int = 1,
                 scale_factor: int = 0, limit: int = 0):
    """(min_time, avg_time, max_time, blend) per blend of a query, fastest
    first: the selection print_base_flavors.py makes."""
    sql = ("SELECT MIN(time_ms) AS min_time, AVG(time_ms) AS avg_time, MAX(time_ms) AS max_time, "
           "default_blend AS flavor FROM runs "
           "WHERE query=? AND mode=? AND threads=? AND aggregates_blend='' AND key_check_blend='' "
           "AND result!='TIMEOUT' AND result!='CRASH' AND scale_factor=? "
           "GROUP BY default_blend ORDER BY avg_time ASC")
    if limit > 0:
        sql += f" LIMIT {int(limit)}"
    return conn.execute(sql, (query, mode, threads, scale_factor)).fetchall()
