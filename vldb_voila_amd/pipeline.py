"""Builder for vh_pipeline_desc: scan -> select -> join* -> group-by.

Mirrors the relalg operators the pipeline is made of (relalg.hpp:98-165):
Scan/Select become a table with conjunctive predicates, HashJoin adds build
payload columns to the stream, HashAggr(keys, sum/count) ends it.
Column data are plain addresses (host for the oracle, device for the GPU).
"""
from __future__ import annotations

from dataclasses import dataclass, field

from . import abi


@dataclass
class Column:
    name: str
    vtype: int
    addr: int
    vmin: int
    vmax: int


@dataclass
class Table:
    name: str
    rows: int
    columns: list = field(default_factory=list)

    def add(self, name, vtype, addr, vmin, vmax):
        self.columns.append(Column(name, vtype, int(addr), int(vmin), int(vmax)))
        return self

    def index(self, name):
        for i, c in enumerate(self.columns):
            if c.name == name:
                return i
        raise KeyError(f"{self.name}: no column {name!r}")


def table_from_arrays(name, arrays: dict, addrs: dict | None = None, bounds: dict | None = None) -> Table:
    """Table over numpy arrays; `addrs` overrides the data address per column
    (device copies of the same arrays); `bounds` gives a column's [min, max]
    metadata (as an importer records it) instead of the data's."""
    rows = None
    t = Table(name, 0)
    for cname, arr in arrays.items():
        if rows is None:
            rows = len(arr)
        assert len(arr) == rows, "ragged table"
        addr = addrs[cname] if addrs is not None else arr.ctypes.data
        if bounds is not None and cname in bounds:
            vmin, vmax = bounds[cname]
        else:
            vmin = int(arr.min()) if len(arr) else 0
            vmax = int(arr.max()) if len(arr) else 0
        t.add(cname, abi.vh_type_of(arr.dtype), addr, vmin, vmax)
    t.rows = rows or 0
    return t


class Expr:
    def __init__(self, op, a=None, b=None, value=0):
        self.op, self.a, self.b, self.value = op, a, b, value

    def __add__(self, o): return Expr(abi.VH_E_ADD, self, _e(o))
    def __radd__(self, o): return Expr(abi.VH_E_ADD, _e(o), self)
    def __sub__(self, o): return Expr(abi.VH_E_SUB, self, _e(o))
    def __rsub__(self, o): return Expr(abi.VH_E_SUB, _e(o), self)
    def __mul__(self, o): return Expr(abi.VH_E_MUL, self, _e(o))
    def __rmul__(self, o): return Expr(abi.VH_E_MUL, _e(o), self)


def col(name):
    return Expr(abi.VH_E_COL, a=name)


def const(v):
    return Expr(abi.VH_E_CONST, value=int(v))


def _e(x):
    return x if isinstance(x, Expr) else const(x)


class Pipeline:
    def __init__(self, probe: Table):
        self.probe = probe
        self.preds = []
        self.joins = []
        self.keys = []
        self.aggrs = []
        self.vnames = [c.name for c in probe.columns]

    def select(self, colname, cmp, value):
        self.preds.append((self.probe.index(colname), abi.CMP_NAMES[cmp], int(value)))
        return self

    def join(self, build: Table, build_key, probe_key, payload=(), variant=abi.VH_JOINN,
             build_preds=(), names=None):
        """build_key / probe_key: a column name, or a list of them for a
        multi-column key (matched pairwise)."""
        names = list(names) if names is not None else [f"{build.name}.{p}" for p in payload]
        bks = [build_key] if isinstance(build_key, str) else list(build_key)
        pks = [probe_key] if isinstance(probe_key, str) else list(probe_key)
        assert len(bks) == len(pks) and 1 <= len(bks) <= abi.VH_MAX_JOIN_KEYS
        self.joins.append(dict(build=build, build_keys=[build.index(k) for k in bks],
                               probe_keys=[self.vindex(k) for k in pks],
                               payload=[build.index(p) for p in payload], variant=variant,
                               preds=[(build.index(c), abi.CMP_NAMES[o], int(v))
                                      for (c, o, v) in build_preds]))
        self.vnames.extend(names)
        return self

    def vindex(self, name):
        return self.vnames.index(name)

    def groupby(self, keys, aggrs):
        """aggrs: list of ("sum", Expr) or ("count",)."""
        self.keys = [self.vindex(k) for k in keys]
        self.aggrs = list(aggrs)
        return self

    # ------------------------------------------------------------------
    @staticmethod
    def _fill_table(td, t: Table, preds):
        td.rows = t.rows
        td.ncols = len(t.columns)
        assert td.ncols <= abi.VH_MAX_COLS
        for i, c in enumerate(t.columns):
            td.cols[i].type = c.vtype
            td.cols[i].data = c.addr
            td.cols[i].min = c.vmin
            td.cols[i].max = c.vmax
        td.npreds = len(preds)
        assert td.npreds <= abi.VH_MAX_PREDS
        for i, (ci, op, v) in enumerate(preds):
            td.preds[i].cmp = op
            td.preds[i].col = ci
            td.preds[i].value = v

    def desc(self) -> abi.vh_pipeline_desc:
        d = abi.vh_pipeline_desc()
        self._fill_table(d.probe, self.probe, self.preds)
        d.njoins = len(self.joins)
        for j, jd in enumerate(self.joins):
            self._fill_table(d.joins[j].build, jd["build"], jd["preds"])
            d.joins[j].build_key = jd["build_keys"][0]
            d.joins[j].probe_key = jd["probe_keys"][0]
            d.joins[j].nkeys = len(jd["build_keys"])
            for k in range(1, len(jd["build_keys"])):
                d.joins[j].build_keys[k - 1] = jd["build_keys"][k]
                d.joins[j].probe_keys[k - 1] = jd["probe_keys"][k]
            d.joins[j].variant = jd["variant"]
            d.joins[j].npayload = len(jd["payload"])
            assert d.joins[j].npayload <= abi.VH_MAX_PAYLOAD
            for p, ci in enumerate(jd["payload"]):
                d.joins[j].payload[p] = ci
        d.nkeys = len(self.keys)
        for k, v in enumerate(self.keys):
            d.keys[k] = v
        nodes = []
        memo = {}

        def emit(e: Expr):
            if id(e) in memo:
                return memo[id(e)]
            if e.op == abi.VH_E_COL:
                n = (abi.VH_E_COL, self.vindex(e.a), 0, 0)
            elif e.op == abi.VH_E_CONST:
                n = (abi.VH_E_CONST, 0, 0, e.value)
            else:
                a = emit(e.a)
                b = emit(e.b)
                n = (e.op, a, b, 0)
            nodes.append(n)
            memo[id(e)] = len(nodes) - 1
            return memo[id(e)]

        d.naggrs = len(self.aggrs)
        for i, ag in enumerate(self.aggrs):
            if ag[0] == "sum":
                d.aggrs[i].fn = abi.VH_A_SUM
                d.aggrs[i].expr = emit(ag[1])
            elif ag[0] == "count":
                d.aggrs[i].fn = abi.VH_A_COUNT
                d.aggrs[i].expr = -1
            else:
                raise ValueError(ag[0])
        assert len(nodes) <= abi.VH_MAX_EXPR
        d.nexpr = len(nodes)
        for i, (op, a, b, v) in enumerate(nodes):
            d.expr[i].op = op
            d.expr[i].a = a
            d.expr[i].b = b
            d.expr[i].value = v
        return d
