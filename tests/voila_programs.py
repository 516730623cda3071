"""VOILA programs as the reference's relalg translator writes them, in the text
form vh_program_lower reads (test infrastructure).

A plan is nested tuples of the relalg operators (relalg.hpp:12-181):
  ("scan", table, [cols])                 ("select", pred, child)
  ("project", [(name, expr) | col], child)
  ("hashjoin", "join01"|"joinn", left, [lkeys], right, [rkeys], [rpayl])
  ("hashaggr", "hash"|"global", child, [keys], [("sum", expr) | ("count",)])
expressions: column names, integers, (op, a, b) with op in + - * and the
comparisons lt le gt ge eq ne, ("and", a, b).

Translator.program(plan) restates RelOpTranslator's statement sequences
(relalg_translator.cpp:202-1115): the scan loop (:202-272), select (:343-378),
project (:275-340), the hash join's materialize / build / probe pipelines
(:812-1115) and the hash aggregation with its flush, read and
re-aggregation pipelines (:381-809).  relalg_text(plan) gives the same plan
as the S-expression vh_plan_lower reads."""


def f(name, args, pred="_"):
    return f"(f {name} ({' '.join(args)}) {pred})"


def ref(n):
    return f"(ref {n})"


def const(v):
    return f"(const {v})"


def tget(i):
    return f("tget", ["lolearg", const(i)])


def assign(var, expr, pred="_"):
    return f"(assign {var} {expr} {pred})"


def emit(expr, pred="_"):
    return f"(emit {expr} {pred})"


def loop(pred, *stmts):
    return f"(loop {pred} {' '.join(stmts)})"


def wrap(pred, *stmts):
    return f"(wrap {pred} {' '.join(stmts)})"


def effect(expr):
    return f"(effect {expr})"


def dead(var):
    return f"(meta vardead {var})"


FUN = {"+": "add", "-": "sub", "*": "mul"}


class Translator:
    def __init__(self):
        self.data = []
        self.pipes = []
        self.cur = []
        self.lole = 0
        self.uniq = 0

    def name(self, op, stage=None):
        self.lole += 1
        n = f"lole_{self.lole}_{op}"
        return f"{n}_{stage}" if stage else n

    def unique(self, prefix):
        self.uniq += 1
        return f"{prefix}{self.uniq}"

    def lolepop(self, name, stmts):
        self.cur.append(f"(lolepop {name} {' '.join(stmts)})")

    def new_pipeline(self):
        self.pipes.append("(pipeline " + " ".join(self.cur) + ")")
        self.cur = []

    # expressions over the flow (ExprTranslator, :20-120)
    def expr(self, e, flow, pred):
        if isinstance(e, int):
            return const(e)
        if isinstance(e, str):
            return tget(flow.index(e))
        op = e[0]
        args = [self.expr(a, flow, pred) for a in e[1:]]
        return f(FUN.get(op, op), args, pred)

    def visit(self, op):
        """Appends op's lolepops to the current pipeline; returns the flow
        (the names of the tuple's columns)."""
        kind = op[0]
        if kind == "scan":
            _, table, cols = op
            colx = [f("scan", [ref(f"{table}.{c}"), ref("pos")]) for c in cols]
            stmts = [
                assign("morsel", f("scan_morsel", [ref(table)])),
                assign("valid_morsel", f("selvalid", [ref("morsel")])),
                loop(ref("valid_morsel"),
                     assign("pos", f("scan_pos", [ref("morsel")])),
                     assign("valid_pos", f("selvalid", [ref("pos")])),
                     loop(ref("valid_pos"),
                          emit(f("tappend", colx)), "(meta refillinflow)",
                          assign("pos", f("scan_pos", [ref("morsel")])),
                          assign("valid_pos", f("selvalid", [ref("pos")]))),
                     dead("pos"), dead("valid_pos"),
                     assign("morsel", f("scan_morsel", [ref(table)])),
                     assign("valid_morsel", f("selvalid", [ref("morsel")]))),
                dead("morsel"), dead("valid_morsel"), "(done)"]
            self.lolepop(self.name("scan"), stmts)
            self.data.append(f"(basetable {table} {table} " +
                             " ".join(f"(col {c} {c} value)" for c in cols) + ")")
            return list(cols)
        if kind == "select":
            _, pred, child = op
            flow = self.visit(child)
            p = f("seltrue", [self.expr(pred, flow, "lolepred")], "lolepred")
            cols = [tget(i) for i in range(len(flow))]
            stmts = [f"(blend lolepred {assign('bv_pred', p, p)})",
                     emit(f("tappend", cols, ref("bv_pred")), ref("bv_pred")), dead("bv_pred")]
            self.lolepop(self.name("select"), stmts)
            return flow
        if kind == "project":
            _, projs, child = op
            flow = self.visit(child)
            cols, names = [], []
            for p in projs:
                if isinstance(p, tuple):
                    names.append(p[0])
                    cols.append(self.expr(p[1], flow, "lolepred"))
                else:
                    names.append(p)
                    cols.append(self.expr(p, flow, "lolepred"))
            self.lolepop(self.name("project"), [emit(f("tappend", cols, "lolepred"), "lolepred")])
            return names
        if kind == "hashjoin":
            return self.hashjoin(*op[1:])
        if kind == "hashaggr":
            return self.hashaggr(*op[1:])
        raise ValueError(kind)

    def hashjoin(self, variant, left, lkeys, right, rkeys, rpayl):
        ht = self.unique("join_ht")
        # materialize (:828-907)
        rflow = self.visit(right)
        stmts = [assign("wpos", f("write_pos", [ref(ht), "lolepred"], "lolepred"), "lolepred")]
        cols, hashx, rcols = [], None, []
        for i, c in enumerate(list(rkeys) + list(rpayl)):
            tc = f"col{i}"
            is_key = i < len(rkeys)
            x = self.expr(c, rflow, "lolepred")
            if is_key:
                hashx = f("rehash", [hashx, x], "lolepred") if hashx else f("hash", [x], "lolepred")
            stmts.append(effect(f("write", [ref(f"{ht}.{tc}"), ref("wpos"), x], "lolepred")))
            cols.append(f"(col {tc} {tc} {'key' if is_key else 'value'})")
            rcols.append(f"{ht}.{tc}")
        hc = f"hash{len(cols)}"
        cols.append(f"(col {hc} {hc} hash)")
        stmts += [effect(f("write", [ref(f"{ht}.{hc}"), ref("wpos"), hashx], "lolepred")), dead("wpos")]
        self.lolepop(self.name("hashjoin", "materialize"), [f"(blend lolepred {' '.join(stmts)})"])
        self.new_pipeline()
        self.data.append(f"(table {ht} hashtable (readafterwrite) {' '.join(cols)})")
        # build (:914-928)
        self.lolepop(self.name("hashjoin", "build"), [effect(f("bucket_build", [ref(ht)])), "(done)"])
        self.new_pipeline()
        # probe (:933-1111)
        flow = self.visit(left)
        hashx, check = None, None
        out = [tget(i) for i in range(len(flow))]
        gathers = []
        for i, k in enumerate(lkeys):
            x = self.expr(k, flow, "lolepred")
            hashx = f("rehash", [hashx, x], "lolepred") if hashx else f("hash", [x], "lolepred")
            c = f("check", [ref(rcols[i]), ref("bucket"), x], ref("active"))
            check = f("and", [check, c], ref("active")) if check else c
            gathers.append(x)
        for i in range(len(rkeys), len(rkeys) + len(rpayl)):
            gathers.append(f("gather", [ref(rcols[i]), ref("bucket")], ref("hit")))
        single = [assign("active", f("selfalse", [ref("match")], ref("active")), ref("active"))] \
            if variant == "join01" else []
        stmts = [
            f"(blend lolepred {assign('bucket', f('bucket_lookup', [ref(ht), hashx], 'lolepred'), 'lolepred')} "
            f"{assign('active', f('selfalse', [f('eq', [const(0), ref('bucket')], 'lolepred')], 'lolepred'), 'lolepred')})",
            loop(ref("active"),
                 f"(blend {ref('active')} {assign('match', check, ref('active'))})",
                 assign("hit", f("seltrue", [ref("match")], ref("active")), ref("active")),
                 emit(f("tappend", out + gathers, ref("hit")), ref("hit")),
                 dead("hit"),
                 wrap(ref("active"), *single),
                 dead("match"),
                 assign("bucket", f("bucket_next", [ref(ht), ref("bucket")], ref("active")), ref("active")),
                 assign("active", f("selfalse", [f("eq", [ref("bucket"), const(0)], ref("active"))],
                                    ref("active")), ref("active"))),
            dead("active"), dead("bucket")]
        self.lolepop(self.name("hashjoin", "probe"), stmts)
        return flow + list(rkeys) + list(rpayl)

    def hashaggr(self, variant, child, keys, aggrs):
        flow = self.visit(child)
        out_keys, out_aggr = self.aggregation(variant, flow, keys, aggrs, reaggr=False)
        # re-aggregation of the flushed partitions (:776-808): counts become sums
        re = [("sum", a) for a in out_aggr]
        return self.aggregation(variant, out_keys + out_aggr, out_keys, re, reaggr=True)[0]

    def aggregation(self, variant, flow, keys, aggrs, reaggr):
        ht = self.unique("aggr_ht")
        cols, stmts, aggr_stmts = [], [], []
        key_cols, visible = [], []
        found = ref("found")
        count_col = None
        global_ = variant == "global"

        def add(a, vis, pred):
            nonlocal count_col
            n = len([c for c in cols if "aggr_" in c])
            short = f"aggr_{n}"
            full = f"{ht}.{short}"
            cols.append(f"(col {short} {short} value)")
            if a[0] == "sum":
                x = self.expr(a[1], flow, pred)
                s = effect(f("aggr_gsum", [ref(full), x], pred)) if global_ else \
                    effect(f("aggr_sum", [ref(full), ref("bucket"), x], pred))
            else:
                count_col = full
                s = effect(f("aggr_gcount", [ref(full)], pred)) if global_ else \
                    effect(f("aggr_count", [ref(full), ref("bucket"), ref("bucket")], pred))
            if vis:
                visible.append(full)
            aggr_stmts.append(s)

        if global_:
            cols.append("(col hash_0 hash_0 hash)")
            for a in aggrs:
                add(a, True, "lolepred")
            if count_col is None:
                add(("count",), False, "lolepred")
            stmts += aggr_stmts
        else:
            hashx, check, scatters = None, None, []
            for i, k in enumerate(keys):
                short = f"key_{i}"
                full = f"{ht}.{short}"
                cols.append(f"(col {short} {short} key)")
                key_cols.append(full)
                x = self.expr(k, flow, found)
                hashx = f("rehash", [hashx, x], "lolepred") if hashx else f("hash", [x], "lolepred")
                c = f("check", [ref(full), ref("bucket"), x], ref("hit"))
                check = f("and", [check, c], ref("hit")) if check else c
                scatters.append(effect(f("scatter", [ref(full), ref("new_pos"), x], ref("can_scatter"))))
            for a in aggrs:
                add(a, True, found)
            if count_col is None:
                add(("count",), False, found)
            hc = f"hash_{len(keys)}"
            cols.append(f"(col {hc} {hc} hash)")
            scatters.append(effect(f("scatter", [ref(f"{ht}.{hc}"), ref("new_pos"), ref("hash")],
                                     ref("can_scatter"))))
            inner = loop(ref("hit"),
                         assign("equal", check, ref("hit")),
                         assign("found", f("seltrue", [ref("equal")], ref("hit")), ref("hit")),
                         wrap(found, *aggr_stmts), dead("found"),
                         assign("hit", f("selfalse", [ref("equal")], ref("hit")), ref("hit")), dead("equal"),
                         assign("bucket", f("bucket_next", [ref(ht), ref("bucket")], ref("hit")), ref("hit")),
                         assign("empty", f("eq", [ref("bucket"), const(0)], ref("hit")), ref("hit")),
                         assign("miss", f("selunion", [ref("miss"), f("seltrue", [ref("empty")], ref("hit"))])),
                         assign("hit", f("selfalse", [ref("empty")], ref("hit")), ref("hit")), dead("empty"))
            outer = loop(ref("miss"),
                         assign("bucket", f("bucket_lookup", [ref(ht), ref("hash")], ref("miss")), ref("miss")),
                         assign("empty", f("eq", [ref("bucket"), const(0)], ref("miss")), ref("miss")),
                         assign("hit", f("selfalse", [ref("empty")], ref("miss")), ref("miss")),
                         assign("miss", f("seltrue", [ref("empty")], ref("miss")), ref("miss")), dead("empty"),
                         inner, dead("bucket"), dead("hit"),
                         assign("new_pos", f("bucket_insert", [ref(ht), ref("hash")], ref("miss")), ref("miss")),
                         assign("can_scatter", f("selfalse", [f("eq", [ref("new_pos"), const(0)], ref("miss"))],
                                                 ref("miss")), ref("miss")),
                         wrap(ref("can_scatter"), *scatters), dead("new_pos"), dead("can_scatter"))
            stmts += [assign("hash", hashx, "lolepred"), assign("miss", "lolepred", "lolepred"),
                      "(meta fsmexclusive begin)", outer, "(meta fsmexclusive end)", dead("hash"), dead("miss")]
        flags = "(threadlocal)" if reaggr else "(threadlocal flushtomaster)"
        self.data.append(f"(table {ht} hashtable {flags} {' '.join(cols)})")
        self.lolepop(self.name("hashaggr", "build"), stmts)
        self.new_pipeline()
        if not reaggr:
            self.lolepop(self.name("hashaggr", "flush"), [effect(f("bucket_flush", [ref(ht)])), "(done)"])
            self.new_pipeline()
        cnt = f("read", [ref(count_col), ref("pos")])
        pred = f("seltrue", [f("gt", [cnt, const(0)])])
        outs = [f("read", [ref(c), ref("pos")]) for c in key_cols + visible]
        self.lolepop(self.name("hashaggr", "read"), [
            assign("morsel", f("read_morsel", [ref(ht)])),
            assign("valid_morsel", f("selvalid", [ref("morsel")])),
            loop(ref("valid_morsel"),
                 assign("pos", f("read_pos", [ref("morsel")])),
                 assign("valid_pos", f("selvalid", [ref("pos")])),
                 loop(ref("valid_pos"), emit(f("tappend", outs, pred), pred), "(meta refillinflow)",
                      assign("pos", f("read_pos", [ref("morsel")])),
                      assign("valid_pos", f("selvalid", [ref("pos")]))),
                 dead("pos"), dead("valid_pos"),
                 assign("morsel", f("read_morsel", [ref(ht)])),
                 assign("valid_morsel", f("selvalid", [ref("morsel")]))),
            dead("morsel"), dead("valid_morsel"), "(done)"])
        return key_cols, visible

    def program(self, plan):
        self.visit(plan)
        if self.cur:
            self.new_pipeline()
        return "(program " + " ".join(self.data) + " " + " ".join(self.pipes) + ")"


def program(plan):
    return Translator().program(plan)


# ---- the same plans as relalg S-expressions (vh_plan_lower) -------------------

def _rexpr(e):
    if isinstance(e, (int, str)):
        return str(e)
    op = e[0]
    names = {"lt": "<", "le": "<=", "gt": ">", "ge": ">=", "eq": "=", "ne": "!="}
    return f"({names.get(op, op)} {' '.join(_rexpr(a) for a in e[1:])})"


def relalg_text(plan):
    kind = plan[0]
    if kind == "scan":
        return f"(scan {plan[1]} {' '.join(plan[2])})"
    if kind == "select":
        return f"(select {_rexpr(plan[1])} {relalg_text(plan[2])})"
    if kind == "project":
        parts = []
        for p in plan[1]:
            if isinstance(p, tuple):
                parts.append(f"(assign {p[0]} {_rexpr(p[1])})")
        return f"(project {' '.join(parts)} {relalg_text(plan[2])})"
    if kind == "hashjoin":
        _, var, left, lk, right, rk, rp = plan
        return (f"(hashjoin {var} {relalg_text(left)} (keys {' '.join(lk)}) (payl) {relalg_text(right)} "
                f"(keys {' '.join(rk)}) (payl {' '.join(rp)}))")
    if kind == "hashaggr":
        _, var, child, keys, aggrs = plan
        ag = " ".join("(count)" if a[0] == "count" else f"(sum {_rexpr(a[1])})" for a in aggrs)
        return f"(hashaggr {var} (keys {' '.join(keys)}) (aggrs {ag}) {relalg_text(child)})"
    raise ValueError(kind)
