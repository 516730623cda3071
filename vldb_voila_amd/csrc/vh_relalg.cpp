// Relational front end of the gfx950 backend ("Flavor::Hip").
//
// Plans use the reference's relational operators (relalg.hpp:12-181): Scan,
// Project, Select, HashAggr{Hash, Global}, HashJoin{Join01, JoinN} with the
// expressions Const, ColId, Fun, Assign.  RelOpTranslator turns such a plan
// into VOILA pipelines (relalg_translator.cpp:181-1124); here the same plan
// shapes are recognised and lowered onto the fused kernels:
//   Scan -> Select*           conjunctive col-vs-const predicates of the probe
//   HashJoin (build = Select*(Scan))  materialise + bucket_build + probe loop
//   Project (Assign)          named expressions used by the aggregates
//   HashAggr                  the group-by loop, flush and re-aggregation
// The text form of a plan is an S-expression (see include/voila_hip.h).
#include <cctype>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "../../include/voila_hip.h"
#include "vh_engine.hpp"

using namespace vh;

#include "vh_sexpr.hpp"

using vh::Sx;
using vh::SxP;
using vh::Parser;
using vh::CatalogTable;

namespace {

// ---- lowering --------------------------------------------------------------

struct Lowering {
	vh_db* db;
	vh_pipeline_desc d;
	std::map<std::string, int> vcol;          // column name -> virtual column
	std::map<std::string, SxP> named;         // Project assignments
	std::string err;
	std::string probe_table;
	std::vector<std::string> explain;

	Lowering(vh_db* db_) : db(db_) { memset(&d, 0, sizeof(d)); }

	bool bad(const std::string& e)
	{
		if (err.empty())
			err = e;
		return false;
	}

	static bool is_int(const std::string& a, int64_t& v)
	{
		if (a.empty())
			return false;
		char* end = nullptr;
		long long x = strtoll(a.c_str(), &end, 10);
		if (*end)
			return false;
		v = x;
		return true;
	}

	// column lookup by "table.col" or bare "col"
	bool find_col(const std::string& tname, const std::string& name, vh_column& out, std::string& col)
	{
		auto t = db->tables.find(tname);
		if (t == db->tables.end())
			return bad("unknown table " + tname);
		std::string c = name;
		const size_t dot = name.find('.');
		if (dot != std::string::npos) {
			if (name.substr(0, dot) != tname)
				return bad("column " + name + " is not in table " + tname);
			c = name.substr(dot + 1);
		}
		auto it = t->second.cols.find(c);
		if (it == t->second.cols.end())
			return bad("unknown column " + tname + "." + c);
		out = it->second;
		col = c;
		return true;
	}

	// Scan: (scan table col...) -> a table descriptor
	bool scan(const SxP& n, vh_table_desc& td, std::vector<std::string>& names, std::string& tname)
	{
		if (!n->is_list() || n->list.size() < 3 || n->list[0]->atom != "scan")
			return bad("expected (scan table col ...)");
		tname = n->list[1]->atom;
		auto t = db->tables.find(tname);
		if (t == db->tables.end())
			return bad("unknown table " + tname);
		memset(&td, 0, sizeof(td));
		td.rows = t->second.rows;
		for (size_t i = 2; i < n->list.size(); i++) {
			vh_column c;
			std::string cn;
			if (!find_col(tname, n->list[i]->atom, c, cn))
				return false;
			if (td.ncols >= VH_MAX_COLS)
				return bad("too many scanned columns");
			td.cols[td.ncols++] = c;
			names.push_back(n->list[i]->atom);
		}
		return true;
	}

	static int cmp_of(const std::string& f)
	{
		// relalg_translator.cpp:24-35: <= le, < lt, >= ge, > gt, = eq, != ne
		if (f == "<") return VH_LT;
		if (f == "<=") return VH_LE;
		if (f == ">") return VH_GT;
		if (f == ">=") return VH_GE;
		if (f == "=" || f == "==") return VH_EQ;
		if (f == "!=" || f == "<>") return VH_NE;
		return -1;
	}

	// conjunction of (cmp col const) over a scanned table
	bool predicates(const SxP& e, vh_table_desc& td, const std::vector<std::string>& names)
	{
		if (!e->is_list() || e->list.size() != 3)
			return bad("selection predicates must be (cmp column constant) or (and p q)");
		const std::string f = e->list[0]->atom;
		if (f == "and" || f == "&&")
			return predicates(e->list[1], td, names) && predicates(e->list[2], td, names);
		int c = cmp_of(f);
		if (c < 0)
			return bad("unsupported predicate function " + f);
		SxP a = e->list[1], b = e->list[2];
		int64_t v;
		if (is_int(a->atom, v) && !is_int(b->atom, v)) {
			std::swap(a, b);   // const op col -> col op' const
			static const int flip[6] = {VH_GT, VH_GE, VH_LT, VH_LE, VH_EQ, VH_NE};
			c = flip[c];
		}
		if (a->is_list() || !is_int(b->atom, v))
			return bad("selection predicates compare a column with a constant");
		int col = -1;
		for (size_t i = 0; i < names.size(); i++)
			if (names[i] == a->atom)
				col = (int)i;
		if (col < 0)
			return bad("predicate column " + a->atom + " is not scanned");
		if (td.npreds >= VH_MAX_PREDS)
			return bad("too many predicates");
		td.preds[td.npreds].cmp = c;
		td.preds[td.npreds].col = col;
		td.preds[td.npreds].value = v;
		td.npreds++;
		return true;
	}

	// build side of a join: Select*(Scan)
	bool build_side(const SxP& n, vh_table_desc& td, std::vector<std::string>& names, std::string& tname)
	{
		if (n->is_list() && !n->list.empty() && n->list[0]->atom == "select") {
			if (n->list.size() != 3)
				return bad("expected (select predicate plan)");
			if (!build_side(n->list[2], td, names, tname))
				return false;
			return predicates(n->list[1], td, names);
		}
		return scan(n, td, names, tname);
	}

	// probe chain: Scan, Select, Project, HashJoin (left deep)
	bool probe_chain(const SxP& n)
	{
		if (!n->is_list() || n->list.empty())
			return bad("expected an operator");
		const std::string op = n->list[0]->atom;
		if (op == "scan") {
			std::vector<std::string> names;
			if (!scan(n, d.probe, names, probe_table))
				return false;
			for (size_t i = 0; i < names.size(); i++)
				vcol[names[i]] = (int)i;
			probe_names = names;
			explain.push_back("scan " + probe_table + ": morsel / pos loop over the base columns (relalg_translator.cpp:202-272)");
			return true;
		}
		if (op == "select") {
			if (n->list.size() != 3)
				return bad("expected (select predicate plan)");
			if (!probe_chain(n->list[2]))
				return false;
			if (d.njoins > 0)
				return bad("a selection above a join is not lowered (push it below the join)");
			explain.push_back("select: vec_<cmp> + vec_seltrue per conjunct (relalg_translator.cpp:343-378)");
			return predicates(n->list[1], d.probe, probe_names);
		}
		if (op == "project") {
			if (n->list.size() < 2)
				return bad("expected (project expr... plan)");
			if (!probe_chain(n->list.back()))
				return false;
			for (size_t i = 1; i + 1 < n->list.size(); i++) {
				const SxP& e = n->list[i];
				if (e->is_list() && !e->list.empty() && e->list[0]->atom == "assign") {
					if (e->list.size() != 3)
						return bad("expected (assign name expr)");
					named[e->list[1]->atom] = e->list[2];
				}
			}
			explain.push_back("project: map primitives (vec_add/sub/mul, casts) evaluated in registers");
			return true;
		}
		if (op == "hashjoin") {
			// (hashjoin join01|joinn LEFT (keys k) (payl ...) RIGHT (keys k) (payl p...))
			if (n->list.size() != 8)
				return bad("expected (hashjoin variant left (keys ..) (payl ..) right (keys ..) (payl ..))");
			const std::string var = n->list[1]->atom;
			if (var != "join01" && var != "joinn")
				return bad("join variant must be join01 or joinn");
			if (!probe_chain(n->list[2]))
				return false;
			if (d.njoins >= VH_MAX_JOINS)
				return bad("too many joins");
			vh_join_desc& J = d.joins[d.njoins];
			memset(&J, 0, sizeof(J));
			std::vector<std::string> bnames;
			std::string btable;
			if (!build_side(n->list[5], J.build, bnames, btable))
				return false;
			const SxP& lk = n->list[3];
			const SxP& rk = n->list[6];
			const SxP& rp = n->list[7];
			if (lk->list.size() != rk->list.size() || lk->list.size() < 2 ||
				lk->list.size() > 1 + VH_MAX_JOIN_KEYS)
				return bad("a join needs 1 to 3 key columns on each side, pairwise");
			J.nkeys = (int)lk->list.size() - 1;
			for (int k = 0; k < J.nkeys; k++) {
				auto lit = vcol.find(lk->list[k + 1]->atom);
				if (lit == vcol.end())
					return bad("probe key " + lk->list[k + 1]->atom + " is not available");
				int bk = -1;
				for (size_t i = 0; i < bnames.size(); i++)
					if (bnames[i] == rk->list[k + 1]->atom)
						bk = (int)i;
				if (bk < 0)
					return bad("build key " + rk->list[k + 1]->atom + " is not scanned");
				if (k == 0) {
					J.probe_key = lit->second;
					J.build_key = bk;
				} else {
					J.probe_keys[k - 1] = lit->second;
					J.build_keys[k - 1] = bk;
				}
			}
			J.variant = var == "join01" ? VH_JOIN01 : VH_JOINN;
			int nv = (int)d.probe.ncols;
			for (int j = 0; j < d.njoins; j++)
				nv += d.joins[j].npayload;
			for (size_t i = 1; i < rp->list.size(); i++) {
				int bc = -1;
				for (size_t k = 0; k < bnames.size(); k++)
					if (bnames[k] == rp->list[i]->atom)
						bc = (int)k;
				if (bc < 0)
					return bad("build payload " + rp->list[i]->atom + " is not scanned");
				if (J.npayload >= VH_MAX_PAYLOAD)
					return bad("too many payload columns");
				J.payload[J.npayload++] = bc;
				vcol[rp->list[i]->atom] = nv++;
			}
			d.njoins++;
			explain.push_back("hashjoin " + var + " with " + btable +
				": materialise + bucket_build, probe loop bucket_lookup / check / gather / bucket_next "
				"(relalg_translator.cpp:812-1115)");
			return true;
		}
		return bad("operator " + op + " is not lowered inside a pipeline");
	}

	std::vector<std::string> probe_names;

	int expr(const SxP& e, int depth = 0)
	{
		if (depth > 32) {
			bad("expression too deep");
			return -1;
		}
		if (d.nexpr >= VH_MAX_EXPR) {
			bad("expression too large");
			return -1;
		}
		if (!e->is_list()) {
			int64_t v;
			vh_expr n;
			memset(&n, 0, sizeof(n));
			if (is_int(e->atom, v)) {
				n.op = VH_E_CONST;
				n.value = v;
			} else {
				auto nm = named.find(e->atom);
				if (nm != named.end())
					return expr(nm->second, depth + 1);
				auto it = vcol.find(e->atom);
				if (it == vcol.end()) {
					bad("unknown column " + e->atom);
					return -1;
				}
				n.op = VH_E_COL;
				n.a = it->second;
			}
			d.expr[d.nexpr] = n;
			return d.nexpr++;
		}
		if (e->list.size() != 3) {
			bad("arithmetic is binary: (+|-|* a b)");
			return -1;
		}
		const std::string f = e->list[0]->atom;
		int op = f == "+" ? VH_E_ADD : f == "-" ? VH_E_SUB : f == "*" ? VH_E_MUL : -1;
		if (op < 0) {
			bad("unsupported function " + f);
			return -1;
		}
		int a = expr(e->list[1], depth + 1);
		int b = a >= 0 ? expr(e->list[2], depth + 1) : -1;
		if (a < 0 || b < 0 || d.nexpr >= VH_MAX_EXPR) {
			bad("expression too large");
			return -1;
		}
		vh_expr n;
		memset(&n, 0, sizeof(n));
		n.op = op;
		n.a = a;
		n.b = b;
		d.expr[d.nexpr] = n;
		return d.nexpr++;
	}

	bool root(const SxP& n)
	{
		// (hashaggr hash|global (keys k...) (aggrs (sum e) (count) ...) PLAN)
		if (!n->is_list() || n->list.size() != 5 || n->list[0]->atom != "hashaggr")
			return bad("the plan root must be (hashaggr hash|global (keys ..) (aggrs ..) plan)");
		const std::string var = n->list[1]->atom;
		if (var != "hash" && var != "global")
			return bad("hashaggr variant must be hash or global");
		if (!probe_chain(n->list[4]))
			return false;
		const SxP& keys = n->list[2];
		if (var == "global" && keys->list.size() > 1)
			return bad("a global aggregate has no keys");
		for (size_t i = 1; i < keys->list.size(); i++) {
			auto it = vcol.find(keys->list[i]->atom);
			if (it == vcol.end())
				return bad("group key " + keys->list[i]->atom + " is not a column of the pipeline");
			if (d.nkeys >= VH_MAX_KEYS)
				return bad("too many group keys");
			d.keys[d.nkeys++] = it->second;
		}
		const SxP& aggrs = n->list[3];
		for (size_t i = 1; i < aggrs->list.size(); i++) {
			const SxP& a = aggrs->list[i];
			if (!a->is_list() || a->list.empty())
				return bad("aggregates are (sum expr) or (count)");
			if (d.naggrs >= VH_MAX_AGGRS)
				return bad("too many aggregates");
			vh_aggr_desc& A = d.aggrs[d.naggrs];
			if (a->list[0]->atom == "count" && a->list.size() == 1) {
				A.fn = VH_A_COUNT;
				A.expr = -1;
			} else if (a->list[0]->atom == "sum" && a->list.size() == 2) {
				A.fn = VH_A_SUM;
				A.expr = expr(a->list[1]);
				if (A.expr < 0)
					return false;
			} else {
				return bad("aggregates are (sum expr) or (count)");
			}
			d.naggrs++;
		}
		explain.push_back(std::string("hashaggr ") + var +
			": hash/rehash, bucket_lookup, check, aggr_sum/aggr_count, bucket_insert + scatter, "
			"bucket_flush and re-aggregation (relalg_translator.cpp:381-809)");
		return true;
	}
};

int lower_text(vh_db* db, const char* plan, Lowering& L)
{
	if (!db || !plan)
		return fail(VH_ERR_INVALID, "null argument");
	Parser p{plan, {}};
	SxP root = p.parse();
	if (!root)
		return fail(VH_ERR_INVALID, "plan: %s", p.err.c_str());
	p.ws();
	if (*p.s)
		return fail(VH_ERR_INVALID, "plan: trailing text");
	if (!L.root(root))
		return fail(VH_ERR_UNSUPPORTED, "plan: %s", L.err.c_str());
	return VH_OK;
}

} // namespace

extern "C" {

int vh_db_new(vh_db** out)
{
	if (!out)
		return fail(VH_ERR_INVALID, "null argument");
	*out = new (std::nothrow) vh_db();
	return *out ? VH_OK : fail(VH_ERR_NOMEM, "host allocation");
}

int vh_db_free(vh_db* db)
{
	delete db;
	return VH_OK;
}

int vh_db_add_column(vh_db* db, const char* table, const char* column, int32_t type, const void* data,
	int64_t rows, int64_t min, int64_t max)
{
	if (!db || !table || !column || type < VH_U8 || type > VH_I64 || rows < 0 || (rows && !data) || min > max)
		return fail(VH_ERR_INVALID, "vh_db_add_column arguments");
	CatalogTable& t = db->tables[table];
	if (t.rows >= 0 && t.rows != rows)
		return fail(VH_ERR_INVALID, "table %s: column %s has %lld rows, table has %lld", table, column,
			(long long)rows, (long long)t.rows);
	t.rows = rows;
	vh_column c;
	memset(&c, 0, sizeof(c));
	c.type = type;
	c.data = data;
	c.min = min;
	c.max = max;
	t.cols[column] = c;
	return VH_OK;
}

int vh_plan_lower(vh_db* db, const char* plan, vh_pipeline_desc* out)
{
	Lowering L(db);
	int rc = lower_text(db, plan, L);
	if (rc)
		return rc;
	if (out)
		memcpy(out, &L.d, sizeof(L.d));
	return VH_OK;
}

int vh_plan_query_new(vh_query** q, vh_db* db, const char* plan, const vh_config* cfg)
{
	vh_pipeline_desc d;
	int rc = vh_plan_lower(db, plan, &d);
	if (rc)
		return rc;
	return vh_query_new(q, &d, cfg);
}

int vh_plan_explain(vh_db* db, const char* plan, char* buf, size_t cap, size_t* len)
{
	if (!len)
		return fail(VH_ERR_INVALID, "null argument");
	Lowering L(db);
	int rc = lower_text(db, plan, L);
	if (rc)
		return rc;
	std::string s;
	for (const std::string& line : L.explain)
		s += line + "\n";
	s += "=> one fused kernel per pipeline: select, probe every join, aggregate\n";
	*len = s.size();
	if (buf) {
		if (cap <= s.size())
			return fail(VH_ERR_CAPACITY, "explain needs %zu bytes", s.size() + 1);
		memcpy(buf, s.c_str(), s.size() + 1);
	}
	return VH_OK;
}

} // extern "C"
