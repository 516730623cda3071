// VOILA programs as input: the Flavor::Hip backend's consumer of the typed
// program the relalg translator produces.
//
// The reference's code generators walk a `Program` (voila.hpp:142-290):
// pipelines of lolepops, each a list of statements over expressions, plus
// the data structures they use.  Program, Pipeline, Lolepop, Statement,
// Expression and DataStructure are restated here with the fields the
// lowering reads; programs arrive as text (the S-expression form below),
// since the C ABI passes no C++ objects.  lower_pipeline recognises the
// statement sequences RelOpTranslator emits (relalg_translator.cpp:202-1115)
// and maps a query's probe pipeline, with the build pipelines of its joins,
// onto one vh_pipeline_desc for the fused kernels:
//   scan lolepop         scan_morsel / scan_pos loops emitting
//                        tappend(scan(T.col, pos) ...)         (:202-272)
//   select lolepop       bv_pred = seltrue(cmp ...), emit | bv_pred (:343-378)
//   project lolepop      emit tappend(expressions)              (:275-340)
//   hashjoin materialize write_pos, write(J.colK, wpos, ...)    (:828-907)
//   hashjoin build       bucket_build(J)                        (:914-928)
//   hashjoin probe       bucket_lookup(J, hash/rehash(keys)), check, gather,
//                        emit, [Join01: active = selfalse(match)], bucket_next
//                                                                (:933-1111)
//   hashaggr build       hash, bucket_lookup / check / bucket_insert loop,
//                        aggr_sum / aggr_count (or aggr_gsum / aggr_gcount),
//                        then flush, read and re-aggregation    (:381-809)
//
// Text form (one expression or statement per list):
//   (program DATA... PIPELINE...)
//   DATA     (basetable NAME SOURCE (col NAME SOURCE MOD)...)
//            (table NAME hashtable|table (FLAG...) (col NAME SOURCE MOD)...)
//            MOD: value | key | hash; FLAG: threadlocal readafterwrite flushtomaster
//   PIPELINE (pipeline (lolepop NAME STMT...)...)
//   STMT     (assign VAR EXPR PRED) (effect EXPR) (emit EXPR PRED)
//            (loop PRED STMT...) (wrap PRED STMT...) (blend PRED STMT...)
//            (done) (meta KIND [VAR])
//   EXPR     (f NAME (ARG...) PRED) (ref NAME) (const TEXT) lolearg lolepred _
//   PRED     an EXPR, or _ for none
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/voila_hip.h"
#include "vh_engine.hpp"
#include "vh_sexpr.hpp"

using namespace vh;

namespace {

// ---- the program (voila.hpp:142-290) ----------------------------------------

struct Expression;
typedef std::shared_ptr<Expression> ExprPtr;

struct Expression {
	enum Type { Function, Constant, Reference, LoleArg, LolePred, NoPred };
	Type type = NoPred;
	std::string fun;                 // function name, constant text or variable
	std::vector<ExprPtr> args;
	ExprPtr pred;
};

struct Statement;
typedef std::shared_ptr<Statement> StmtPtr;

struct Statement {
	enum Type { Loop, Assignment, EffectExpr, Emit, Done, Wrap, BlendStmt, MetaStmt };
	Type type = Done;
	std::string var;
	ExprPtr expr;
	ExprPtr pred;
	std::vector<StmtPtr> statements;
};

struct Lolepop {
	std::string name;
	std::vector<StmtPtr> statements;
};

struct Pipeline {
	std::vector<Lolepop> lolepops;
};

struct DCol {
	enum Modifier { kValue, kKey, kHash };
	std::string name;
	std::string source;
	Modifier mod = kValue;
};

struct DataStructure {
	enum Type { kTable, kHashTable, kBaseTable };
	static constexpr int kThreadLocal = 1 << 1;
	static constexpr int kReadAfterWrite = 1 << 2;
	static constexpr int kFlushToMaster = 1 << 3;
	std::string name;
	std::string source;
	Type type = kTable;
	int flags = 0;
	std::vector<DCol> cols;
};

struct Program {
	std::vector<Pipeline> pipelines;
	std::vector<DataStructure> data_structures;
};

// ---- reading the text form ------------------------------------------------

struct Reader {
	std::string err;

	bool bad(const std::string& e)
	{
		if (err.empty())
			err = e;
		return false;
	}

	ExprPtr expr(const SxP& n)
	{
		auto e = std::make_shared<Expression>();
		if (!n->is_list()) {
			if (n->atom == "lolearg")
				e->type = Expression::LoleArg;
			else if (n->atom == "lolepred")
				e->type = Expression::LolePred;
			else if (n->atom == "_")
				return nullptr;
			else {
				bad("unknown expression atom " + n->atom);
				return nullptr;
			}
			return e;
		}
		if (n->list.empty() || n->list[0]->is_list()) {
			bad("malformed expression");
			return nullptr;
		}
		const std::string& h = n->list[0]->atom;
		if ((h == "ref" || h == "const") && n->list.size() == 2 && !n->list[1]->is_list()) {
			e->type = h == "ref" ? Expression::Reference : Expression::Constant;
			e->fun = n->list[1]->atom;
			return e;
		}
		if (h == "f" && n->list.size() == 4 && !n->list[1]->is_list() && n->list[2]->is_list()) {
			e->type = Expression::Function;
			e->fun = n->list[1]->atom;
			for (auto& a : n->list[2]->list) {
				auto x = expr(a);
				if (!x) {
					bad("missing argument of " + e->fun);
					return nullptr;
				}
				e->args.push_back(x);
			}
			e->pred = expr(n->list[3]);
			return err.empty() ? e : nullptr;
		}
		bad("unknown expression " + h);
		return nullptr;
	}

	StmtPtr stmt(const SxP& n)
	{
		if (!n->is_list() || n->list.empty() || n->list[0]->is_list()) {
			bad("malformed statement");
			return nullptr;
		}
		auto s = std::make_shared<Statement>();
		const std::string& h = n->list[0]->atom;
		const size_t m = n->list.size();
		if (h == "assign" && m == 4 && !n->list[1]->is_list()) {
			s->type = Statement::Assignment;
			s->var = n->list[1]->atom;
			s->expr = expr(n->list[2]);
			s->pred = expr(n->list[3]);
		} else if (h == "effect" && m == 2) {
			s->type = Statement::EffectExpr;
			s->expr = expr(n->list[1]);
		} else if (h == "emit" && m == 3) {
			s->type = Statement::Emit;
			s->expr = expr(n->list[1]);
			s->pred = expr(n->list[2]);
		} else if ((h == "loop" || h == "wrap" || h == "blend") && m >= 2) {
			s->type = h == "loop" ? Statement::Loop : (h == "wrap" ? Statement::Wrap : Statement::BlendStmt);
			s->pred = expr(n->list[1]);
			for (size_t i = 2; i < m; i++) {
				auto c = stmt(n->list[i]);
				if (!c)
					return nullptr;
				s->statements.push_back(c);
			}
		} else if (h == "done" && m == 1) {
			s->type = Statement::Done;
		} else if (h == "meta" && m >= 2) {
			s->type = Statement::MetaStmt;
			s->var = m > 2 && !n->list[2]->is_list() ? n->list[2]->atom : "";
		} else {
			bad("unknown statement " + h);
			return nullptr;
		}
		return err.empty() ? s : nullptr;
	}

	bool columns(const SxP& n, size_t first, DataStructure& d)
	{
		for (size_t i = first; i < n->list.size(); i++) {
			const SxP& c = n->list[i];
			if (!c->is_list() || c->list.size() != 4 || c->list[0]->atom != "col")
				return bad("malformed column of " + d.name);
			DCol col;
			col.name = c->list[1]->atom;
			col.source = c->list[2]->atom;
			const std::string& m = c->list[3]->atom;
			col.mod = m == "key" ? DCol::kKey : (m == "hash" ? DCol::kHash : DCol::kValue);
			if (m != "key" && m != "hash" && m != "value")
				return bad("column modifier " + m);
			d.cols.push_back(col);
		}
		return true;
	}

	bool program(const SxP& n, Program& p)
	{
		if (!n->is_list() || n->list.empty() || n->list[0]->atom != "program")
			return bad("expected (program ...)");
		for (size_t i = 1; i < n->list.size(); i++) {
			const SxP& x = n->list[i];
			if (!x->is_list() || x->list.empty())
				return bad("malformed program entry");
			const std::string& h = x->list[0]->atom;
			if (h == "basetable" && x->list.size() >= 3) {
				DataStructure d;
				d.type = DataStructure::kBaseTable;
				d.name = x->list[1]->atom;
				d.source = x->list[2]->atom;
				if (!columns(x, 3, d))
					return false;
				p.data_structures.push_back(d);
			} else if (h == "table" && x->list.size() >= 4 && x->list[3]->is_list()) {
				DataStructure d;
				d.name = x->list[1]->atom;
				d.type = x->list[2]->atom == "hashtable" ? DataStructure::kHashTable : DataStructure::kTable;
				for (auto& f : x->list[3]->list) {
					if (f->atom == "threadlocal")
						d.flags |= DataStructure::kThreadLocal;
					else if (f->atom == "readafterwrite")
						d.flags |= DataStructure::kReadAfterWrite;
					else if (f->atom == "flushtomaster")
						d.flags |= DataStructure::kFlushToMaster;
					else
						return bad("table flag " + f->atom);
				}
				if (!columns(x, 4, d))
					return false;
				p.data_structures.push_back(d);
			} else if (h == "pipeline") {
				Pipeline pipe;
				for (size_t k = 1; k < x->list.size(); k++) {
					const SxP& l = x->list[k];
					if (!l->is_list() || l->list.size() < 2 || l->list[0]->atom != "lolepop")
						return bad("expected (lolepop NAME STMT...)");
					Lolepop lp;
					lp.name = l->list[1]->atom;
					for (size_t q = 2; q < l->list.size(); q++) {
						auto s = stmt(l->list[q]);
						if (!s)
							return bad("in lolepop " + lp.name);
						lp.statements.push_back(s);
					}
					pipe.lolepops.push_back(lp);
				}
				p.pipelines.push_back(pipe);
			} else {
				return bad("unknown program entry " + h);
			}
		}
		return true;
	}
};

// ---- walking statements ---------------------------------------------------

bool is_fun(const ExprPtr& e, const char* name)
{
	return e && e->type == Expression::Function && e->fun == name;
}

// Every function expression in statements (depth first, in order).
void collect_funs(const std::vector<StmtPtr>& ss, std::vector<ExprPtr>& out)
{
	std::vector<ExprPtr> stack;
	auto walk = [&](const ExprPtr& e, auto& self) -> void {
		if (!e)
			return;
		if (e->type == Expression::Function)
			out.push_back(e);
		for (auto& a : e->args)
			self(a, self);
		self(e->pred, self);
	};
	for (auto& s : ss) {
		walk(s->expr, walk);
		walk(s->pred, walk);
		collect_funs(s->statements, out);
	}
}

std::vector<ExprPtr> funs_of(const Lolepop& l)
{
	std::vector<ExprPtr> v;
	collect_funs(l.statements, v);
	return v;
}

bool has_fun(const Lolepop& l, const char* name)
{
	for (auto& f : funs_of(l))
		if (f->fun == name)
			return true;
	return false;
}

// The assignments of a lolepop, anywhere in it: var -> expression (the last
// one wins, as the statement order overwrites).
void collect_assigns(const std::vector<StmtPtr>& ss, std::map<std::string, ExprPtr>& out)
{
	for (auto& s : ss) {
		if (s->type == Statement::Assignment)
			out[s->var] = s->expr;
		collect_assigns(s->statements, out);
	}
}

// The tappend of the lolepop's (last) emit.
ExprPtr emitted(const std::vector<StmtPtr>& ss)
{
	ExprPtr r;
	for (auto& s : ss) {
		if (s->type == Statement::Emit && is_fun(s->expr, "tappend"))
			r = s->expr;
		if (auto x = emitted(s->statements))
			r = x;
	}
	return r;
}

// ---- lowering ----------------------------------------------------------------

// A value of the tuple flow between lolepops.
struct Val {
	enum Kind { Col, Node, Const } kind = Col;
	int vcol = -1;            // Col: virtual column of the descriptor
	ExprPtr expr;             // Node: expression over earlier flow values
	std::vector<Val> env;     // Node: the flow the expression's tget indexes
	int64_t c = 0;            // Const
	std::string table, column;   // Col of a base table (build sides)
};

struct Lower {
	vh_db* db;
	const Program& prog;
	vh_pipeline_desc d;
	std::string err;
	std::map<std::string, const DataStructure*> ds;

	Lower(vh_db* db_, const Program& p) : db(db_), prog(p) { memset(&d, 0, sizeof(d)); }

	bool bad(const std::string& e)
	{
		if (err.empty())
			err = e;
		return false;
	}

	static bool parse_int(const std::string& s, int64_t& v)
	{
		char* end = nullptr;
		v = strtoll(s.c_str(), &end, 10);
		return !s.empty() && end && *end == 0;
	}

	// a base table's scan lolepop: the flow of its columns (table.col refs)
	bool scan_flow(const Lolepop& l, std::string& table, std::vector<Val>& flow)
	{
		table.clear();
		for (auto& f : funs_of(l))
			if (f->fun == "scan_morsel" && f->args.size() == 1 && f->args[0]->type == Expression::Reference)
				table = f->args[0]->fun;
		auto it = ds.find(table);
		if (table.empty() || it == ds.end() || it->second->type != DataStructure::kBaseTable)
			return bad("scan lolepop " + l.name + " does not scan a base table");
		const std::string src = it->second->source;
		auto t = db->tables.find(src);
		if (t == db->tables.end())
			return bad("table " + src + " is not in the catalog");
		ExprPtr tap = emitted(l.statements);
		if (!tap)
			return bad("scan lolepop " + l.name + " emits no tuple");
		flow.clear();
		for (auto& a : tap->args) {
			if (!is_fun(a, "scan") || a->args.empty() || a->args[0]->type != Expression::Reference)
				return bad("scan lolepop " + l.name + ": expected scan(T.col, pos)");
			std::string ref = a->args[0]->fun;
			const size_t dot = ref.find('.');
			const std::string col = dot == std::string::npos ? ref : ref.substr(dot + 1);
			if (!t->second.cols.count(col))
				return bad("column " + col + " is not in the catalog's " + src);
			Val v;
			v.kind = Val::Col;
			v.table = src;
			v.column = col;
			flow.push_back(v);
		}
		table = src;
		return true;
	}

	// value of expression e over flow (tget(lolearg, i) = flow[i])
	bool value(const ExprPtr& e, const std::vector<Val>& flow, Val& out)
	{
		if (!e)
			return bad("missing expression");
		if (e->type == Expression::Constant) {
			out.kind = Val::Const;
			if (!parse_int(e->fun, out.c))
				return bad("constant " + e->fun + " is not an integer");
			return true;
		}
		if (is_fun(e, "tget")) {
			int64_t i = -1;
			if (e->args.size() != 2 || e->args[1]->type != Expression::Constant || !parse_int(e->args[1]->fun, i) ||
				i < 0 || i >= (int64_t)flow.size())
				return bad("tget outside the tuple");
			out = flow[(size_t)i];
			return true;
		}
		if (e->type == Expression::Function &&
			(e->fun == "add" || e->fun == "sub" || e->fun == "mul") && e->args.size() == 2) {
			out.kind = Val::Node;
			out.expr = e;
			out.env = flow;
			return true;
		}
		return bad("expression " + e->fun + " is not lowered");
	}

	// conjunctive column-vs-constant predicates of a seltrue(...) argument
	bool preds_of(const ExprPtr& e, const std::vector<Val>& flow, std::vector<vh_pred>& out,
		std::vector<std::string>* cols_out)
	{
		if (is_fun(e, "and") && e->args.size() == 2)
			return preds_of(e->args[0], flow, out, cols_out) && preds_of(e->args[1], flow, out, cols_out);
		static const std::map<std::string, int> cmps = {{"lt", VH_LT}, {"le", VH_LE}, {"gt", VH_GT}, {"ge", VH_GE},
			{"eq", VH_EQ}, {"ne", VH_NE}};
		static const std::map<int, int> flip = {{VH_LT, VH_GT}, {VH_LE, VH_GE}, {VH_GT, VH_LT}, {VH_GE, VH_LE},
			{VH_EQ, VH_EQ}, {VH_NE, VH_NE}};
		if (!e || e->type != Expression::Function || !cmps.count(e->fun) || e->args.size() != 2)
			return bad("predicate " + (e ? e->fun : std::string("?")) + " is not lowered");
		Val a, b;
		if (!value(e->args[0], flow, a) || !value(e->args[1], flow, b))
			return false;
		int cmp = cmps.at(e->fun);
		if (a.kind == Val::Const && b.kind == Val::Col) {
			std::swap(a, b);
			cmp = flip.at(cmp);
		}
		if (a.kind != Val::Col || b.kind != Val::Const)
			return bad("predicates compare a column with a constant");
		vh_pred p;
		memset(&p, 0, sizeof(p));
		p.cmp = cmp;
		p.value = b.c;
		p.col = a.vcol;
		if (cols_out)
			cols_out->push_back(a.column);
		out.push_back(p);
		return true;
	}

	// the predicate of a select lolepop: bv_pred = seltrue(...)
	bool select_preds(const Lolepop& l, const std::vector<Val>& flow, std::vector<vh_pred>& out,
		std::vector<std::string>* cols_out)
	{
		std::map<std::string, ExprPtr> as;
		collect_assigns(l.statements, as);
		for (auto& kv : as)
			if (is_fun(kv.second, "seltrue") && kv.second->args.size() == 1)
				return preds_of(kv.second->args[0], flow, out, cols_out);
		return bad("select lolepop " + l.name + " has no seltrue predicate");
	}

	static bool is_select(const Lolepop& l)
	{
		std::map<std::string, ExprPtr> as;
		collect_assigns(l.statements, as);
		for (auto& kv : as)
			if (is_fun(kv.second, "seltrue"))
				return !has_fun(l, "bucket_lookup") && !has_fun(l, "bucket_insert");
		return false;
	}

	// ---- joins: build pipelines ---------------------------------------------

	struct JoinBuild {
		std::string table;                 // catalog table of the build side
		std::vector<vh_pred> preds;
		std::vector<std::string> pred_cols;
		std::map<std::string, std::string> col_of;   // J.colK -> build column
		std::vector<std::string> keys;     // J columns with kKey, in order
		std::vector<std::string> pays;     // J columns with kValue, in order
	};
	std::map<std::string, JoinBuild> builds;

	bool find_build(const std::string& J)
	{
		if (builds.count(J))
			return true;
		const DataStructure* t = ds.at(J);
		JoinBuild B;
		for (auto& c : t->cols) {
			const std::string full = J + "." + c.name;
			if (c.mod == DCol::kKey)
				B.keys.push_back(full);
			else if (c.mod == DCol::kValue)
				B.pays.push_back(full);
		}
		for (auto& pipe : prog.pipelines) {
			int mat = -1;
			for (size_t i = 0; i < pipe.lolepops.size(); i++)
				for (auto& f : funs_of(pipe.lolepops[i]))
					if (f->fun == "write" && f->args.size() == 3 && f->args[0]->type == Expression::Reference &&
						f->args[0]->fun.compare(0, J.size() + 1, J + ".") == 0)
						mat = (int)i;
			if (mat < 0)
				continue;
			std::vector<Val> flow;
			if (mat < 1 || !scan_flow(pipe.lolepops[0], B.table, flow))
				return bad("build side of " + J + " does not start with a scan");
			for (size_t i = 0; i < flow.size(); i++)
				flow[i].vcol = (int)i;
			for (int i = 1; i < mat; i++) {
				if (!is_select(pipe.lolepops[i]))
					return bad("build side of " + J + ": only selections between scan and materialize");
				if (!select_preds(pipe.lolepops[i], flow, B.preds, &B.pred_cols))
					return false;
			}
			for (auto& f : funs_of(pipe.lolepops[mat])) {
				if (f->fun != "write" || f->args.size() != 3 || f->args[0]->type != Expression::Reference)
					continue;
				const std::string& dst = f->args[0]->fun;
				if (is_fun(f->args[2], "hash") || is_fun(f->args[2], "rehash"))
					continue;   // the hash column
				Val v;
				if (!value(f->args[2], flow, v))
					return false;
				if (v.kind != Val::Col)
					return bad("build side of " + J + " writes an expression, not a column");
				B.col_of[dst] = v.column;
			}
			for (auto& k : B.keys)
				if (!B.col_of.count(k))
					return bad("key column " + k + " of " + J + " is never written");
			for (auto& k : B.pays)
				if (!B.col_of.count(k))
					return bad("payload column " + k + " of " + J + " is never written");
			builds[J] = B;
			return true;
		}
		return bad("no pipeline materializes " + J);
	}

	// ---- the probe pipeline ---------------------------------------------------

	bool fill_table(vh_table_desc& td, const std::string& table, std::vector<std::string>& names)
	{
		const CatalogTable& t = db->tables.at(table);
		td.rows = t.rows;
		td.ncols = 0;
		for (auto& n : names) {
			if (td.ncols >= VH_MAX_COLS)
				return bad("table " + table + ": more than 8 columns");
			td.cols[td.ncols++] = t.cols.at(n);
		}
		return true;
	}

	static int col_index(std::vector<std::string>& names, const std::string& c)
	{
		for (size_t i = 0; i < names.size(); i++)
			if (names[i] == c)
				return (int)i;
		names.push_back(c);
		return (int)names.size() - 1;
	}

	int emit_node(const Val& v)
	{
		if (d.nexpr >= VH_MAX_EXPR) {
			bad("expression too large");
			return -1;
		}
		vh_expr& e = d.expr[d.nexpr];
		memset(&e, 0, sizeof(e));
		if (v.kind == Val::Const) {
			e.op = VH_E_CONST;
			e.value = v.c;
			return d.nexpr++;
		}
		if (v.kind == Val::Col) {
			e.op = VH_E_COL;
			e.a = v.vcol;
			return d.nexpr++;
		}
		Val a, b;
		if (!value(v.expr->args[0], v.env, a) || !value(v.expr->args[1], v.env, b))
			return -1;
		const int ia = emit_node(a);
		const int ib = ia < 0 ? -1 : emit_node(b);
		if (ib < 0 || d.nexpr >= VH_MAX_EXPR) {
			bad("expression too large");
			return -1;
		}
		vh_expr& n = d.expr[d.nexpr];
		memset(&n, 0, sizeof(n));
		n.op = v.expr->fun == "add" ? VH_E_ADD : (v.expr->fun == "sub" ? VH_E_SUB : VH_E_MUL);
		n.a = ia;
		n.b = ib;
		return d.nexpr++;
	}

	// keys of a hash(k0) / rehash(hash(k0), k1) ... chain, in order
	bool hash_keys(const ExprPtr& h, std::vector<ExprPtr>& keys)
	{
		if (is_fun(h, "hash") && h->args.size() == 1) {
			keys.push_back(h->args[0]);
			return true;
		}
		if (is_fun(h, "rehash") && h->args.size() == 2) {
			if (!hash_keys(h->args[0], keys))
				return false;
			keys.push_back(h->args[1]);
			return true;
		}
		return bad("expected a hash / rehash chain of keys");
	}

	bool lower_program()
	{
		for (auto& x : prog.data_structures)
			ds[x.name] = &x;
		// the probe pipeline: the one whose last lolepop is the first-level
		// aggregation (a table flushed to the master table)
		const Pipeline* probe = nullptr;
		std::string aggr;
		for (auto& pipe : prog.pipelines) {
			if (pipe.lolepops.empty())
				continue;
			const Lolepop& l = pipe.lolepops.back();
			for (auto& f : funs_of(l)) {
				if ((f->fun == "bucket_insert" || f->fun == "aggr_gsum" || f->fun == "aggr_gcount") && !f->args.empty() &&
					f->args[0]->type == Expression::Reference) {
					std::string t = f->args[0]->fun;
					const size_t dot = t.find('.');
					if (f->fun != "bucket_insert" && dot != std::string::npos)
						t = t.substr(0, dot);
					auto it = ds.find(t);
					if (it != ds.end() && (it->second->flags & DataStructure::kFlushToMaster)) {
						probe = &pipe;
						aggr = t;
					}
				}
			}
		}
		if (!probe)
			return bad("no pipeline ends in a first-level aggregation (a hash table flushed to the master)");
		std::string table;
		std::vector<Val> flow;
		if (!scan_flow(probe->lolepops[0], table, flow))
			return false;
		std::vector<std::string> pnames;
		for (auto& v : flow)
			v.vcol = col_index(pnames, v.column);
		nprobe = (int)pnames.size();   // payload columns are numbered after these
		std::vector<vh_pred> preds;
		std::vector<std::string> probe_pred_cols;
		std::vector<std::vector<std::string>> payload_names;   // per join: build column names
		const size_t L = probe->lolepops.size();
		for (size_t i = 1; i + 1 < L; i++) {
			const Lolepop& l = probe->lolepops[i];
			if (has_fun(l, "bucket_lookup")) {
				if (!join_lolepop(l, flow))
					return false;
			} else if (is_select(l)) {
				if (d.njoins)
					return bad("a selection after a join is not lowered (only selections of the probe table)");
				if (!select_preds(l, flow, preds, &probe_pred_cols))
					return false;
			} else {
				// project: the emitted expressions become the flow
				ExprPtr tap = emitted(l.statements);
				if (!tap)
					return bad("lolepop " + l.name + " is not recognised");
				std::map<std::string, ExprPtr> as;
				collect_assigns(l.statements, as);
				std::vector<Val> next;
				for (auto& a : tap->args) {
					ExprPtr e = a;
					if (e->type == Expression::Reference && as.count(e->fun))
						e = as[e->fun];   // bv_i variables of a blended projection
					Val v;
					if (!value(e, flow, v))
						return false;
					next.push_back(v);
				}
				flow = next;
			}
		}
		// probe table columns: as scanned (predicate columns included)
		if (!fill_table(d.probe, table, pnames))
			return false;
		d.probe.npreds = (int)preds.size();
		if (d.probe.npreds > VH_MAX_PREDS)
			return bad("more than 8 predicates");
		for (size_t k = 0; k < preds.size(); k++)
			d.probe.preds[k] = preds[k];
		// virtual column numbers were assigned in order: probe columns then
		// each join's payloads; the probe table keeps every scanned column
		if ((int)pnames.size() != d.probe.ncols)
			return bad("internal: probe columns");
		return aggregation(probe->lolepops.back(), aggr, flow);
	}

	// probe lolepop of join J: keys from bucket_lookup(J, hash(...)), the
	// variant from the single-match statement, the emitted tuple's layout
	bool join_lolepop(const Lolepop& l, std::vector<Val>& flow)
	{
		ExprPtr lookup;
		for (auto& f : funs_of(l))
			if (f->fun == "bucket_lookup")
				lookup = f;
		if (!lookup || lookup->args.size() != 2 || lookup->args[0]->type != Expression::Reference)
			return bad("probe lolepop " + l.name + ": malformed bucket_lookup");
		const std::string J = lookup->args[0]->fun;
		if (!ds.count(J) || !(ds.at(J)->flags & DataStructure::kReadAfterWrite))
			return bad("probe lolepop " + l.name + " looks up " + J + ", not a join table");
		if (!find_build(J))
			return false;
		const JoinBuild& B = builds.at(J);
		std::vector<ExprPtr> keys;
		if (!hash_keys(lookup->args[1], keys))
			return false;
		if (keys.size() != B.keys.size() || keys.empty() || keys.size() > VH_MAX_JOIN_KEYS)
			return bad("join " + J + ": key count differs between probe and build");
		if (d.njoins >= VH_MAX_JOINS)
			return bad("more than 4 joins");
		vh_join_desc& jd = d.joins[d.njoins];
		memset(&jd, 0, sizeof(jd));
		// Join01: the loop's single-match statement active = selfalse(match)
		jd.variant = VH_JOINN;
		std::map<std::string, ExprPtr> as;
		collect_assigns(l.statements, as);
		std::vector<ExprPtr> funs = funs_of(l);
		for (auto& f : funs)
			if (f->fun == "selfalse" && f->args.size() == 1 && f->args[0]->type == Expression::Reference &&
				f->args[0]->fun == "match")
				jd.variant = VH_JOIN01;
		// build table: its key and payload columns, then predicate columns
		std::vector<std::string> bnames;
		for (auto& k : B.keys)
			col_index(bnames, B.col_of.at(k));
		for (auto& p : B.pays)
			col_index(bnames, B.col_of.at(p));
		for (auto& c : B.pred_cols)
			col_index(bnames, c);
		if (!fill_table(jd.build, B.table, bnames))
			return false;
		jd.build.npreds = (int)B.preds.size();
		for (size_t k = 0; k < B.preds.size(); k++) {
			jd.build.preds[k] = B.preds[k];
			jd.build.preds[k].col = col_index(bnames, B.pred_cols[k]);
		}
		jd.nkeys = (int)keys.size();
		for (size_t k = 0; k < keys.size(); k++) {
			Val v;
			if (!value(keys[k], flow, v))
				return false;
			if (v.kind != Val::Col)
				return bad("join " + J + ": a probe key is an expression, not a column");
			const int bk = col_index(bnames, B.col_of.at(B.keys[k]));
			if (k == 0) {
				jd.probe_key = v.vcol;
				jd.build_key = bk;
			} else {
				jd.probe_keys[k - 1] = v.vcol;
				jd.build_keys[k - 1] = bk;
			}
		}
		if ((int)B.pays.size() > VH_MAX_PAYLOAD)
			return bad("join " + J + ": more than 4 payload columns");
		jd.npayload = (int)B.pays.size();
		for (size_t k = 0; k < B.pays.size(); k++)
			jd.payload[k] = col_index(bnames, B.col_of.at(B.pays[k]));
		// the emitted tuple: the input's columns, then the table's columns
		// in table order (keys are the probe keys' values, payloads gathers)
		ExprPtr tap = emitted(l.statements);
		if (!tap)
			return bad("probe lolepop " + l.name + " emits no tuple");
		const int first_pay = vcols_before_join();
		std::vector<Val> next;
		for (auto& a : tap->args) {
			ExprPtr e = a;
			if (e->type == Expression::Reference && as.count(e->fun))
				e = as[e->fun];   // bv_i variables of a blended reconstruction
			if (is_fun(e, "gather") && e->args.size() == 2 && e->args[0]->type == Expression::Reference) {
				const std::string& c = e->args[0]->fun;
				int k = -1;
				for (size_t q = 0; q < B.pays.size(); q++)
					if (B.pays[q] == c)
						k = (int)q;
				if (k < 0)
					return bad("gather of " + c + ", not a payload of " + J);
				Val v;
				v.kind = Val::Col;
				v.vcol = first_pay + k;
				v.table = B.table;
				v.column = B.col_of.at(c);
				next.push_back(v);
			} else {
				Val v;
				if (!value(e, flow, v))
					return false;
				next.push_back(v);
			}
		}
		d.njoins++;
		flow = next;
		return true;
	}

	// the first virtual column of the next join's payloads
	int vcols_before_join()
	{
		int v = nprobe;
		for (int j = 0; j < d.njoins; j++)
			v += d.joins[j].npayload;
		return v;
	}

	int nprobe = 0;

	// the aggregation lolepop: group keys and aggregates
	bool aggregation(const Lolepop& l, const std::string& aggr, const std::vector<Val>& flow)
	{
		std::map<std::string, ExprPtr> as;
		collect_assigns(l.statements, as);
		std::vector<ExprPtr> keys;
		if (as.count("hash")) {
			if (!hash_keys(as["hash"], keys))
				return false;
		}
		if ((int)keys.size() > VH_MAX_KEYS)
			return bad("more than 4 group keys");
		d.nkeys = (int)keys.size();
		for (size_t k = 0; k < keys.size(); k++) {
			Val v;
			if (!value(keys[k], flow, v))
				return false;
			if (v.kind != Val::Col)
				return bad("a group key is an expression, not a column");
			d.keys[k] = v.vcol;
		}
		// the visible aggregates: what the read lolepop of the table emits
		// (the implicit count of :465-472 is not among them)
		std::vector<std::string> visible;
		for (auto& pipe : prog.pipelines)
			for (auto& lp : pipe.lolepops) {
				bool reads = false;
				for (auto& f : funs_of(lp))
					if (f->fun == "read_morsel" && f->args.size() == 1 && f->args[0]->fun == aggr)
						reads = true;
				if (!reads)
					continue;
				ExprPtr tap = emitted(lp.statements);
				if (!tap)
					continue;
				for (auto& a : tap->args)
					if (is_fun(a, "read") && !a->args.empty() && a->args[0]->type == Expression::Reference)
						visible.push_back(a->args[0]->fun);
			}
		std::map<std::string, ExprPtr> aggr_fun;   // aggregate column -> its aggr_* call
		for (auto& f : funs_of(l))
			if ((f->fun == "aggr_sum" || f->fun == "aggr_count" || f->fun == "aggr_gsum" || f->fun == "aggr_gcount") &&
				!f->args.empty() && f->args[0]->type == Expression::Reference)
				aggr_fun[f->args[0]->fun] = f;
		int na = 0;
		for (auto& c : visible) {
			if (!aggr_fun.count(c))
				continue;   // a key column
			const ExprPtr& f = aggr_fun[c];
			if (na >= VH_MAX_AGGRS)
				return bad("more than 8 aggregates");
			vh_aggr_desc& a = d.aggrs[na++];
			if (f->fun == "aggr_count" || f->fun == "aggr_gcount") {
				a.fn = VH_A_COUNT;
				a.expr = -1;
				continue;
			}
			a.fn = VH_A_SUM;
			Val v;
			if (!value(f->args.back(), flow, v))
				return false;
			if ((a.expr = emit_node(v)) < 0)
				return false;
		}
		d.naggrs = na;
		if (!na && !d.nkeys)
			return bad("the aggregation has neither keys nor aggregates");
		return true;
	}
};

int lower_text(vh_db* db, const char* text, vh_pipeline_desc* out, std::string* err)
{
	Parser ps{text};
	SxP root = ps.parse();
	if (!root)
		return fail(VH_ERR_INVALID, "program: %s", ps.err.c_str());
	Program prog;
	Reader rd;
	if (!rd.program(root, prog))
		return fail(VH_ERR_INVALID, "program: %s", rd.err.c_str());
	Lower lw(db, prog);
	if (!lw.lower_program()) {
		if (err)
			*err = lw.err;
		return fail(VH_ERR_UNSUPPORTED, "program: %s", lw.err.c_str());
	}
	*out = lw.d;
	return VH_OK;
}

} // namespace

extern "C" {

int vh_program_lower(vh_db* db, const char* program, vh_pipeline_desc* out)
{
	if (!db || !program || !out)
		return fail(VH_ERR_INVALID, "vh_program_lower arguments");
	return lower_text(db, program, out, nullptr);
}

int vh_program_query_new(vh_query** q, vh_db* db, const char* program, const vh_config* cfg)
{
	if (!q || !db || !program)
		return fail(VH_ERR_INVALID, "vh_program_query_new arguments");
	vh_pipeline_desc d;
	int rc = lower_text(db, program, &d, nullptr);
	if (rc)
		return rc;
	return vh_query_new(q, &d, cfg);
}

} // extern "C"
