"""Python binding of the relational front end (vh_db_*, vh_plan_*).

Plans are the reference's relalg operators (relalg.hpp:12-181) written as
S-expressions, e.g. the filter + group-by workload:

    (hashaggr hash (keys flag status)
      (aggrs (sum qty) (sum price) (sum (* price (- 100 disc))) (count))
      (select (<= date 2451058) (scan lineitem flag status qty price disc date)))
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import abi
from .runtime import Query, check, lib


def _bind():
    L = lib()
    L.vh_db_new.argtypes = [C.POINTER(C.c_void_p)]
    L.vh_db_free.argtypes = [C.c_void_p]
    L.vh_db_add_column.argtypes = [C.c_void_p, C.c_char_p, C.c_char_p, C.c_int32, C.c_void_p, C.c_int64,
                                   C.c_int64, C.c_int64]
    L.vh_plan_lower.argtypes = [C.c_void_p, C.c_char_p, C.c_void_p]
    L.vh_plan_query_new.argtypes = [C.POINTER(C.c_void_p), C.c_void_p, C.c_char_p, C.c_void_p]
    L.vh_plan_explain.argtypes = [C.c_void_p, C.c_char_p, C.c_char_p, C.c_size_t, C.POINTER(C.c_size_t)]
    L.vh_program_lower.argtypes = [C.c_void_p, C.c_char_p, C.c_void_p]
    return L


class Database:
    """runtime::Database analog: named tables of typed columns with min/max."""

    def __init__(self):
        self.L = _bind()
        p = C.c_void_p()
        check(self.L.vh_db_new(C.byref(p)))
        self.ptr = p.value

    def add_table(self, name, arrays: dict, addrs: dict | None = None, bounds: dict | None = None):
        """bounds: a column's [min, max] as an importer recorded it (default:
        the data's)."""
        for col, arr in arrays.items():
            addr = addrs[col] if addrs is not None else arr.ctypes.data
            if bounds is not None and col in bounds:
                lo, hi = bounds[col]
            else:
                lo = int(arr.min()) if len(arr) else 0
                hi = int(arr.max()) if len(arr) else 0
            check(self.L.vh_db_add_column(self.ptr, name.encode(), col.encode(), abi.vh_type_of(arr.dtype),
                                          addr, len(arr), lo, hi))
        return self

    def lower(self, plan: str) -> abi.vh_pipeline_desc:
        d = abi.vh_pipeline_desc()
        check(self.L.vh_plan_lower(self.ptr, plan.encode(), C.byref(d)))
        return d

    def lower_program(self, program: str) -> abi.vh_pipeline_desc:
        """A VOILA program (the text form of include/voila_hip.h) lowered to
        the pipeline descriptor its probe pipeline and joins make."""
        d = abi.vh_pipeline_desc()
        check(self.L.vh_program_lower(self.ptr, program.encode(), C.byref(d)))
        return d

    def query_program(self, program: str, config=None) -> Query:
        return Query(self.lower_program(program), config)

    def explain(self, plan: str) -> str:
        n = C.c_size_t()
        check(self.L.vh_plan_explain(self.ptr, plan.encode(), None, 0, C.byref(n)))
        buf = C.create_string_buffer(n.value + 1)
        check(self.L.vh_plan_explain(self.ptr, plan.encode(), buf, n.value + 1, C.byref(n)))
        return buf.value.decode()

    def query(self, plan: str, config=None) -> Query:
        """Flavor::Hip query over device columns."""
        return Query(self.lower(plan), config)

    def free(self):
        if self.ptr:
            self.L.vh_db_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def filter_groupby_plan(cutoff: int) -> str:
    return f"""(hashaggr hash (keys flag status)
  (aggrs (sum qty) (sum price) (sum (* price (- 100 disc))) (count))
  (select (<= date {cutoff}) (scan lineitem flag status qty price disc date)))"""


def join_aggregate_plan(cutoff: int, variant="joinn") -> str:
    return f"""(hashaggr hash (keys segment) (aggrs (sum price))
  (hashjoin {variant}
    (select (<= date {cutoff}) (scan probe fk date price)) (keys fk) (payl)
    (scan build key segment) (keys key) (payl segment)))"""


def tpch_q1_plan(cutoff: int) -> str:
    """TPC-H Q1 over imported lineitem columns (decimals scaled by 100, so
    1 - l_discount is 100 - l_discount); the averages are the sums over the
    count."""
    return f"""(hashaggr hash (keys l_returnflag l_linestatus)
  (aggrs (sum l_quantity) (sum l_extendedprice) (sum (* l_extendedprice (- 100 l_discount)))
         (sum (* (* l_extendedprice (- 100 l_discount)) (+ 100 l_tax))) (count))
  (select (<= l_shipdate {cutoff})
    (scan lineitem l_returnflag l_linestatus l_quantity l_extendedprice l_discount l_tax l_shipdate)))"""


def tpch_q6_plan(date_from: int, date_to: int) -> str:
    """TPC-H Q6: discount between 0.05 and 0.07, quantity below 24."""
    return f"""(hashaggr global (keys) (aggrs (sum (* l_extendedprice l_discount)))
  (select (and (>= l_shipdate {date_from}) (and (< l_shipdate {date_to})
          (and (>= l_discount 5) (and (<= l_discount 7) (< l_quantity 2400)))))
    (scan lineitem l_shipdate l_discount l_quantity l_extendedprice)))"""


def star_plan(region: int) -> str:
    return f"""(hashaggr hash (keys c_nation s_nation d_year)
  (aggrs (sum (- f_revenue f_cost)) (count))
  (hashjoin joinn
    (hashjoin joinn
      (hashjoin joinn
        (hashjoin joinn
          (scan fact f_date f_cust f_supp f_part f_revenue f_cost) (keys f_cust) (payl)
          (select (= c_region {region}) (scan customer c_key c_region c_nation)) (keys c_key) (payl c_nation))
        (keys f_supp) (payl)
        (select (= s_region {region}) (scan supplier s_key s_region s_nation)) (keys s_key) (payl s_nation))
      (keys f_date) (payl)
      (scan date d_key d_year) (keys d_key) (payl d_year))
    (keys f_part) (payl)
    (scan part p_key) (keys p_key) (payl)))"""
