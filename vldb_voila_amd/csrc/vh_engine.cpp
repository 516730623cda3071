// Host side of the gfx950 backend: typing, lowering of a vh_pipeline_desc
// into the fused-kernel plan, device buffers, the query lifecycle and the
// QueryResult text.  Everything here is plain host C++ over the HIP runtime.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/voila_hip.h"
#include "vh_batch.hpp"
#include "vh_engine.hpp"
#include "vh_jit.hpp"
#include "vh_kernels.hpp"

namespace vh {

// ------------------------------------------------------------------------
// errors
// ------------------------------------------------------------------------

static thread_local std::string g_last_error;

int fail(int code, const char* fmt, ...)
{
	char buf[512];
	va_list ap;
	va_start(ap, fmt);
	vsnprintf(buf, sizeof(buf), fmt, ap);
	va_end(ap);
	g_last_error = buf;
	return code;
}

const char* last_error() { return g_last_error.c_str(); }

#define HIP_TRY(x)                                                                    \
	do {                                                                              \
		hipError_t _e = (x);                                                          \
		if (_e != hipSuccess)                                                         \
			return fail(VH_ERR_HIP, "%s failed: %s", #x, hipGetErrorString(_e));     \
	} while (0)

// ------------------------------------------------------------------------
// typing: restates typing_pass.cpp for the values a pipeline produces
// ------------------------------------------------------------------------

static void type_limits(int t, double& lo, double& hi)
{
	switch (t) {
	case VH_U8: lo = 0; hi = 255.0; break;
	case VH_I8: lo = -128.0; hi = 127.0; break;
	case VH_U16: lo = 0; hi = 65535.0; break;
	case VH_I16: lo = -32768.0; hi = 32767.0; break;
	case VH_U32: lo = 0; hi = 4294967295.0; break;
	case VH_I32: lo = -2147483648.0; hi = 2147483647.0; break;
	case VH_U64: lo = 0; hi = (double)UINT64_MAX; break;
	case VH_I64: lo = (double)INT64_MIN; hi = (double)INT64_MAX; break;
	default: lo = -1.7e38; hi = 1.7e38; break;
	}
}

// type_from_minmax (typing_pass.cpp:9-22)
int type_from_minmax(double mn, double mx)
{
	static const int order[8] = {VH_U8, VH_I8, VH_U16, VH_I16, VH_U32, VH_I32, VH_U64, VH_I64};
	for (int t : order) {
		double lo, hi;
		type_limits(t, lo, hi);
		if (mn >= lo && mx <= hi)
			return t;
	}
	return VH_I128;
}

// scan type: signed -> unsigned when min >= 0 (typing_pass.cpp:309-326)
static int scan_type(int t, int64_t mn)
{
	if (mn < 0)
		return t;
	switch (t) {
	case VH_I8: return VH_U8;
	case VH_I16: return VH_U16;
	case VH_I32: return VH_U32;
	default: return t;
	}
}

static bool type_signed(int t)
{
	return t == VH_I8 || t == VH_I16 || t == VH_I32 || t == VH_I64 || t == VH_I128;
}

// Compute units of the current device (256 on MI355X), looked up once.
static int device_cus()
{
	static int cus = 0;
	if (!cus) {
		int dev = 0, n = 0;
		if (hipGetDevice(&dev) == hipSuccess &&
			hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && n > 0)
			cus = n;
		else
			cus = 256;
	}
	return cus;
}

// GB_PART: groups per partition slice (a power of two whose LDS slice,
// 12 bytes per sum + a 4-byte count per group, fits in 32 KiB)
static int part_shift(int nsums)
{
	int shift = 12;
	while (shift > 6 && ((size_t)1 << shift) * ((size_t)nsums * 12 + 4) > 32 * 1024)
		shift--;
	return shift;
}

// GB_PART: the tuples of a region staged in LDS (the rest go through gtup0)
int part_stage(const Plan& P)
{
	return (int)std::min<i64>(kPartStageBytes / P.grec, P.gregion_rows);
}

// GB_PART slices: the largest the LDS allows (2^part_shift groups), halved
// (down to 64 groups) while half the slice would still give fewer than
// `want` partitions (a workgroup or work item each)
void part_partitions(Plan& P, i64 want)
{
	P.gshift = part_shift(P.nsums);
	while (P.gshift > 6 && (i64)(P.domain >> (P.gshift - 1)) < want)
		P.gshift--;
	P.gparts = (int)((P.domain + (1ull << P.gshift) - 1) >> P.gshift);
}

static bool part_fits(double domain, int nsums)
{
	return nsums <= 4 && domain <= (double)kMaxGroupParts * (double)(1u << part_shift(nsums));
}

static const double kMaxCard = 1099511627776.0;   // QueryConfig::max_card = 2^40

int type_pipeline(const vh_pipeline_desc* d, Typing& ty)
{
	ty.vcols.clear();
	for (int c = 0; c < d->probe.ncols; c++) {
		const vh_column& col = d->probe.cols[c];
		ty.vcols.push_back({scan_type(col.type, col.min), (double)col.min, (double)col.max,
			(double)col.min, (double)col.max});
	}
	for (int j = 0; j < d->njoins; j++) {
		const vh_join_desc& jd = d->joins[j];
		if (jd.probe_key < 0 || jd.probe_key >= (int)ty.vcols.size())
			return fail(VH_ERR_INVALID, "join %d: probe_key %d out of range", j, jd.probe_key);
		if (jd.nkeys < 0 || jd.nkeys > VH_MAX_JOIN_KEYS)
			return fail(VH_ERR_INVALID, "join %d: nkeys %d", j, jd.nkeys);
		for (int k = 1; k < jd.nkeys; k++)
			if (jd.probe_keys[k - 1] < 0 || jd.probe_keys[k - 1] >= (int)ty.vcols.size() ||
				jd.build_keys[k - 1] < 0 || jd.build_keys[k - 1] >= jd.build.ncols)
				return fail(VH_ERR_INVALID, "join %d: key column %d out of range", j, k);
		if (jd.npayload < 0 || jd.npayload > VH_MAX_PAYLOAD)
			return fail(VH_ERR_INVALID, "join %d: npayload %d", j, jd.npayload);
		for (int p = 0; p < jd.npayload; p++) {
			if (jd.payload[p] < 0 || jd.payload[p] >= jd.build.ncols)
				return fail(VH_ERR_INVALID, "join %d: payload column %d", j, jd.payload[p]);
			const vh_column& col = jd.build.cols[jd.payload[p]];
			ty.vcols.push_back({scan_type(col.type, col.min), (double)col.min, (double)col.max,
				(double)col.min, (double)col.max});
		}
	}
	ty.nodes.assign(d->nexpr, TypeProps{});
	for (int n = 0; n < d->nexpr; n++) {
		const vh_expr& e = d->expr[n];
		TypeProps t{};
		switch (e.op) {
		case VH_E_COL:
			if (e.a < 0 || e.a >= (int)ty.vcols.size())
				return fail(VH_ERR_INVALID, "expr %d: column %d", n, e.a);
			t = ty.vcols[e.a];
			break;
		case VH_E_CONST:
			t.dmin = t.dmax = t.emin = t.emax = (double)e.value;
			t.type = type_from_minmax(t.dmin, t.dmax);
			break;
		case VH_E_ADD:
		case VH_E_SUB:
		case VH_E_MUL: {
			if (e.a < 0 || e.a >= n || e.b < 0 || e.b >= n)
				return fail(VH_ERR_INVALID, "expr %d: children must precede the node", n);
			const TypeProps& A = ty.nodes[e.a];
			const TypeProps& B = ty.nodes[e.b];
			double x, y;
			// ARITH2_INFIX range rule (typing_pass.cpp:164-176)
			if (e.op == VH_E_ADD) { x = A.dmin + B.dmin; y = A.dmax + B.dmax; }
			else if (e.op == VH_E_SUB) { x = A.dmin - B.dmax; y = A.dmax - B.dmin; }
			else { x = A.dmin * B.dmin; y = A.dmax * B.dmax; }
			t.dmin = std::min(x, y);
			t.dmax = std::max(x, y);
			t.type = type_from_minmax(t.dmin, t.dmax);
			// exact interval of the true value
			if (e.op == VH_E_ADD) { t.emin = A.emin + B.emin; t.emax = A.emax + B.emax; }
			else if (e.op == VH_E_SUB) { t.emin = A.emin - B.emax; t.emax = A.emax - B.emin; }
			else {
				double c[4] = {A.emin * B.emin, A.emin * B.emax, A.emax * B.emin, A.emax * B.emax};
				t.emin = *std::min_element(c, c + 4);
				t.emax = *std::max_element(c, c + 4);
			}
			break;
		}
		default:
			return fail(VH_ERR_INVALID, "expr %d: op %d", n, e.op);
		}
		ty.nodes[n] = t;
	}
	ty.aggrs.assign(d->naggrs, TypeProps{});
	ty.final_aggrs.assign(d->naggrs, TypeProps{});
	for (int a = 0; a < d->naggrs; a++) {
		const vh_aggr_desc& ad = d->aggrs[a];
		TypeProps t{};
		if (ad.fn == VH_A_SUM) {
			if (ad.expr < 0 || ad.expr >= d->nexpr)
				return fail(VH_ERR_INVALID, "aggregate %d: expr %d", a, ad.expr);
			t.dmin = kMaxCard * ty.nodes[ad.expr].dmin;
			t.dmax = kMaxCard * ty.nodes[ad.expr].dmax;
		} else if (ad.fn == VH_A_COUNT) {
			t.dmin = 0;
			t.dmax = kMaxCard;
		} else {
			return fail(VH_ERR_INVALID, "aggregate %d: fn %d", a, ad.fn);
		}
		t.type = type_from_minmax(t.dmin, t.dmax);
		ty.aggrs[a] = t;
		// re-aggregation (count -> sum) of the flushed partitions
		TypeProps f{};
		f.dmin = kMaxCard * t.dmin;
		f.dmax = kMaxCard * t.dmax;
		f.type = type_from_minmax(f.dmin, f.dmax);
		ty.final_aggrs[a] = f;
	}
	for (int k = 0; k < d->nkeys; k++)
		if (d->keys[k] < 0 || d->keys[k] >= (int)ty.vcols.size())
			return fail(VH_ERR_INVALID, "key %d: column %d", k, d->keys[k]);
	return VH_OK;
}

// ------------------------------------------------------------------------
// lowering
// ------------------------------------------------------------------------

// calc_num_buckets (runtime_struct.cpp:790-814)
static uint64_t calc_num_buckets(uint64_t count)
{
	uint64_t n = std::max<uint64_t>({count, 2048, 2 * 1024});
	int exp = 64 - __builtin_clzll(n);
	if ((double)(1ull << exp) < (double)n / 0.7)
		exp++;
	return 1ull << exp;
}

static int role_of_vcol(const vh_pipeline_desc* d, int v, Role& r)
{
	memset(&r, 0, sizeof(r));
	if (v < d->probe.ncols) {
		const vh_column& c = d->probe.cols[v];
		r.kind = ROLE_COL;
		r.ptr = c.data;
		r.type = c.type;
		return VH_OK;
	}
	int base = d->probe.ncols;
	for (int j = 0; j < d->njoins; j++) {
		if (v < base + d->joins[j].npayload) {
			r.kind = ROLE_PAY;
			r.slot = j;
			r.item = v - base;
			r.type = d->joins[j].build.cols[d->joins[j].payload[v - base]].type;
			return VH_OK;
		}
		base += d->joins[j].npayload;
	}
	return fail(VH_ERR_INVALID, "virtual column %d", v);
}

// Flatten an ADD/SUB tree over <= 2 columns and constants into sa*a + sb*b + c.
static bool linear_term(const vh_pipeline_desc* d, int n, int sign, Term& T, int& ncols)
{
	const vh_expr& e = d->expr[n];
	switch (e.op) {
	case VH_E_CONST:
		T.c += sign * e.value;
		return true;
	case VH_E_COL: {
		Role r;
		if (role_of_vcol(d, e.a, r) != VH_OK)
			return false;
		if (ncols == 0) { T.a = r; T.sa = sign; }
		else if (ncols == 1) { T.b = r; T.sb = sign; }
		else return false;
		ncols++;
		return true;
	}
	case VH_E_ADD:
		return linear_term(d, e.a, sign, T, ncols) && linear_term(d, e.b, sign, T, ncols);
	case VH_E_SUB:
		return linear_term(d, e.a, sign, T, ncols) && linear_term(d, e.b, -sign, T, ncols);
	default:
		return false;
	}
}

static bool product_terms(const vh_pipeline_desc* d, int n, SumArg& s)
{
	const vh_expr& e = d->expr[n];
	if (e.op == VH_E_MUL)
		return product_terms(d, e.a, s) && product_terms(d, e.b, s);
	if (s.nterms >= 3)
		return false;
	Term& T = s.t[s.nterms];
	memset(&T, 0, sizeof(T));
	int ncols = 0;
	if (!linear_term(d, n, 1, T, ncols))
		return false;
	s.nterms++;
	return true;
}

static bool fits_type(int t, double lo, double hi)
{
	double tl, th;
	type_limits(t, tl, th);
	return lo >= tl && hi <= th && t != VH_I128;
}

int lower(const vh_pipeline_desc* d, const vh_config* cfg, Lowered& L)
{
	memset(&L.plan, 0, sizeof(L.plan));
	int rc = type_pipeline(d, L.ty);
	if (rc)
		return rc;
	Plan& P = L.plan;
	P.rows = d->probe.rows;
	if (P.rows < 0)
		return fail(VH_ERR_INVALID, "negative row count");
	if (d->probe.ncols < 0 || d->probe.ncols > VH_MAX_COLS || d->probe.npreds < 0 ||
		d->probe.npreds > VH_MAX_PREDS || d->njoins < 0 || d->njoins > VH_MAX_JOINS ||
		d->nkeys < 0 || d->nkeys > VH_MAX_KEYS || d->naggrs < 0 || d->naggrs > VH_MAX_AGGRS ||
		d->nexpr < 0 || d->nexpr > VH_MAX_EXPR)
		return fail(VH_ERR_INVALID, "descriptor counts out of range");
	for (int c = 0; c < d->probe.ncols; c++)
		if (d->probe.rows > 0 && !d->probe.cols[c].data)
			return fail(VH_ERR_INVALID, "probe column %d has no data", c);

	// selection
	P.npreds = d->probe.npreds;
	for (int p = 0; p < P.npreds; p++) {
		const vh_pred& pr = d->probe.preds[p];
		if (pr.col < 0 || pr.col >= d->probe.ncols || pr.cmp < VH_LT || pr.cmp > VH_NE)
			return fail(VH_ERR_INVALID, "predicate %d", p);
		role_of_vcol(d, pr.col, P.preds[p].r);
		P.preds[p].cmp = pr.cmp;
		P.preds[p].value = pr.value;
	}

	// joins
	P.njoins = d->njoins;
	L.builds.assign(d->njoins, BuildPlan{});
	L.capacity.assign(d->njoins, 0);
	L.dir_words.assign(d->njoins, 0);
	const int layout = cfg ? cfg->join_layout : VH_LAYOUT_AUTO;
	if (layout < VH_LAYOUT_AUTO || layout > VH_LAYOUT_DIRECT)
		return fail(VH_ERR_INVALID, "join_layout %d", layout);
	for (int j = 0; j < d->njoins; j++) {
		const vh_join_desc& jd = d->joins[j];
		BuildPlan& B = L.builds[j];
		memset(&B, 0, sizeof(B));
		if (jd.build_key < 0 || jd.build_key >= jd.build.ncols)
			return fail(VH_ERR_INVALID, "join %d: build_key %d", j, jd.build_key);
		if (jd.variant != VH_JOIN01 && jd.variant != VH_JOINN)
			return fail(VH_ERR_INVALID, "join %d: variant %d", j, jd.variant);
		const vh_column& kc = jd.build.cols[jd.build_key];
		const int nk = std::max(jd.nkeys, 1);
		// a multi-column key: each build column's offset from its minimum in
		// a bit field wide enough for its range; the fields together form
		// the entry's key (at most 32 bits).  A probe value outside the
		// build column's range matches no row.
		int kbits = 0;
		u64 kpacked_max = 0;
		if (nk > 1) {
			for (int c = 0; c < nk; c++) {
				const int bc = c ? jd.build_keys[c - 1] : jd.build_key;
				const vh_column& col = jd.build.cols[bc];
				const double lo = (double)col.min, hi = (double)col.max;
				const double range = hi - lo;
				int bits = 0;
				while (bits < 40 && (double)(1ull << bits) <= range)
					bits++;
				if (kbits + bits > 32)
					return fail(VH_ERR_UNSUPPORTED, "join %d: key columns span more than 32 bits", j);
				kpacked_max |= (u64)range << kbits;
				kbits += bits;
			}
			if (kpacked_max >= 0xFFFFFFFFull)
				return fail(VH_ERR_UNSUPPORTED, "join %d: key columns span more than 32 bits", j);
		} else if ((double)kc.max - (double)kc.min >= 4294967295.0) {
			return fail(VH_ERR_UNSUPPORTED, "join %d: key range exceeds 32 bits", j);
		}
		B.rows = jd.build.rows;
		B.npreds = jd.build.npreds;
		if (B.npreds < 0 || B.npreds > VH_MAX_PREDS)
			return fail(VH_ERR_INVALID, "join %d: npreds", j);
		for (int p = 0; p < B.npreds; p++) {
			const vh_pred& pr = jd.build.preds[p];
			if (pr.col < 0 || pr.col >= jd.build.ncols || pr.cmp < VH_LT || pr.cmp > VH_NE)
				return fail(VH_ERR_INVALID, "join %d: predicate %d", j, p);
			Role& r = B.preds[p].r;
			memset(&r, 0, sizeof(r));
			r.kind = ROLE_COL;
			r.ptr = jd.build.cols[pr.col].data;
			r.type = jd.build.cols[pr.col].type;
			B.preds[p].cmp = pr.cmp;
			B.preds[p].value = pr.value;
		}
		B.key = kc.data;
		B.key_type = kc.type;
		B.kmin = nk > 1 ? 0 : kc.min;
		B.kmax = nk > 1 ? (i64)kpacked_max : kc.max;
		B.nkeys = nk;
		if (nk > 1) {
			int shift = 0;
			for (int c = 0; c < nk; c++) {
				const vh_column& col = jd.build.cols[c ? jd.build_keys[c - 1] : jd.build_key];
				B.keyc[c] = col.data;
				B.keyc_type[c] = col.type;
				B.kminc[c] = col.min;
				B.krangec[c] = (u64)(col.max - col.min);
				B.kshiftc[c] = shift;
				int bits = 0;
				while ((double)(1ull << bits) <= (double)col.max - (double)col.min)
					bits++;
				shift += bits;
			}
		}
		// payload columns: bit fields of one 32-bit word (a single column
		// keeps its whole range below 2^32 - 1)
		B.npay = jd.npayload;
		u64 pmax = 0;
		if (jd.npayload == 1) {
			const vh_column& pc = jd.build.cols[jd.payload[0]];
			if ((double)pc.max - (double)pc.min >= 4294967295.0)
				return fail(VH_ERR_UNSUPPORTED, "join %d: payload range exceeds 32 bits", j);
			B.pay = pc.data;
			B.pay_type = pc.type;
			B.pmin = pc.min;
			B.payc[0] = pc.data;
			B.payc_type[0] = pc.type;
			B.pminc[0] = pc.min;
			pmax = (u64)(pc.max - pc.min);
		} else if (jd.npayload > 1) {
			int shift = 0;
			for (int k = 0; k < jd.npayload; k++) {
				const vh_column& pc = jd.build.cols[jd.payload[k]];
				const double range = (double)pc.max - (double)pc.min;
				int bits = 0;
				while (bits < 40 && (double)(1ull << bits) <= range)
					bits++;
				if (shift + bits > 32)
					return fail(VH_ERR_UNSUPPORTED, "join %d: payload columns span more than 32 bits", j);
				B.payc[k] = pc.data;
				B.payc_type[k] = pc.type;
				B.pminc[k] = pc.min;
				B.pshiftc[k] = shift;
				pmax |= (u64)range << shift;
				shift += bits;
			}
			if (pmax >= 0xFFFFFFFFull)
				return fail(VH_ERR_UNSUPPORTED, "join %d: payload columns span more than 32 bits", j);
			B.pmin = B.pminc[0];
		}
		uint64_t cap = calc_num_buckets((uint64_t)std::max<int64_t>(jd.build.rows, 0));
		if (cfg && cfg->table_shift > 0)
			cap <<= cfg->table_shift;
		if (cfg && cfg->table_shift < 0)
			cap = std::max<uint64_t>(cap >> (-cfg->table_shift), 2048);
		while ((double)cap < (double)jd.build.rows / 0.7)   // never above 70% load
			cap <<= 1;
		L.capacity[j] = cap;
		B.mask = cap - 1;

		// direct layout: an entry per key of [kmin, kmax], payload - pmin in
		// the fewest bytes that keep all-ones free as "no row"
		const double krange = (double)B.kmax - (double)B.kmin + 1;
		const double prange = (double)pmax;
		const bool dense = krange <= 4.0 * (double)std::max<int64_t>(jd.build.rows, 0) + 65536.0;
		const int pcfg = cfg ? cfg->part_shift : 0;
		if (pcfg < -1 || pcfg > 30)
			return fail(VH_ERR_INVALID, "part_shift %d", pcfg);
		const uint64_t brows = (uint64_t)std::max<int64_t>(jd.build.rows, 1);
		const uint64_t chunks = (brows + kBuildChunkRows - 1) / kBuildChunkRows;
		if (layout != VH_LAYOUT_HASH && krange <= (double)(1u << 30) && (layout == VH_LAYOUT_DIRECT || dense)) {
			B.dir_w = prange <= 254 ? 1 : (prange <= 65534 ? 2 : 4);
			// whole 16-byte units, so that an LDS copy may read it 16 bytes at a time
			L.dir_words[j] = ((uint64_t)krange * (uint64_t)B.dir_w + 15) / 16 * 2;
			// partitioned build: slices of about 8192 rows (one workgroup
			// each), at most kMaxSliceBytes of LDS and kMaxBuildParts of them
			const uint64_t R = (uint64_t)krange;
			int shift = 10;
			if (pcfg > 0)
				shift = std::max(pcfg, 5);   // the key bitmap takes whole 32-bit words of a slice
			else
				while (shift < 30 && ((uint64_t)1 << shift) * brows < R * 8192)
					shift++;
			while (shift > 5 && ((uint64_t)1 << shift) * (uint64_t)B.dir_w > (uint64_t)kMaxSliceBytes)
				shift--;
			const uint64_t parts = (R + ((uint64_t)1 << shift) - 1) >> shift;
			if (pcfg >= 0 && parts <= (uint64_t)kMaxBuildParts && chunks <= (uint64_t)kMaxBuildChunks) {
				B.pshift = shift;
				B.nparts = (int)parts;
				B.nchunks = jd.build.rows > 0 ? (int)chunks : 0;
			}
		} else if (pcfg >= 0) {
			// hash table: partitioned build when it exceeds one XCD's 4 MiB
			// of L2 (or when asked): blocks of 2^13 slots (64 KiB of LDS),
			// larger when there would be more than kMaxBuildParts of them
			const int cbits = 63 - __builtin_clzll(cap);
			int shift = pcfg > 0 ? std::min(std::max(pcfg, 10), 14) : 13;
			while (shift < 14 && (cap >> shift) > (uint64_t)kMaxBuildParts)
				shift++;
			shift = std::min(shift, cbits);
			if ((pcfg > 0 || cap * sizeof(u64) > ((uint64_t)4 << 20)) && (cap >> shift) <= (uint64_t)kMaxBuildParts &&
				chunks <= (uint64_t)kMaxBuildChunks) {
				B.hashed = 1;
				B.pshift = shift;
				B.nparts = (int)(cap >> shift);
				B.nchunks = jd.build.rows > 0 ? (int)chunks : 0;
			}
		}

		JoinT& J = P.joins[j];
		J.mask = cap - 1;
		J.smask = B.hashed ? ((1ull << B.pshift) - 1) : cap - 1;
		J.hashed = B.hashed;
		if ((rc = role_of_vcol(d, jd.probe_key, J.key)))
			return rc;
		J.kmin = B.kmin;
		J.kmax = B.kmax;
		J.pmin = B.pmin;
		J.dir_w = B.dir_w;
		J.npay = jd.npayload;
		J.nkeys = nk;
		if (nk > 1) {
			for (int c = 0; c < nk; c++) {
				if ((rc = role_of_vcol(d, c ? jd.probe_keys[c - 1] : jd.probe_key, J.keyc[c])))
					return rc;
				J.kminc[c] = B.kminc[c];
				J.krangec[c] = B.krangec[c];
				J.kshiftc[c] = B.kshiftc[c];
			}
		}
		for (int k = 0; k < VH_MAX_PAYLOAD; k++) {
			J.pminc[k] = B.pminc[k];
			J.pshiftc[k] = B.pshiftc[k];
			J.pmaskc[k] = 0xFFFFFFFFu;
		}
		if (jd.npayload > 1) {
			for (int k = 0; k < jd.npayload; k++) {
				const int next = k + 1 < jd.npayload ? B.pshiftc[k + 1] : 32;
				const int bits = next - B.pshiftc[k];
				J.pmaskc[k] = bits >= 32 ? 0xFFFFFFFFu : ((1u << bits) - 1u);
			}
		}
		J.variant = jd.variant;
		J.last = j == d->njoins - 1;
		J.filtered = B.npreds > 0;
	}

	// group keys
	P.nkeys = d->nkeys;
	double domain = 1;
	int bits = 0;
	for (int k = 0; k < d->nkeys; k++) {
		const TypeProps& t = L.ty.vcols[d->keys[k]];
		if ((rc = role_of_vcol(d, d->keys[k], P.keys[k].r)))
			return rc;
		P.keys[k].min = (int64_t)t.dmin;
		double range = t.dmax - t.dmin + 1;
		domain *= range;
		int b = 0;
		while (b < 63 && (double)(1ull << b) < range)
			b++;
		P.keys[k].shift = bits;
		P.keys[k].range = (u64)(t.dmax - t.dmin);
		bits += b;
		L.key_range[k] = (uint64_t)range;
		L.key_bits[k] = b;
	}
	{
		uint64_t mult = 1;
		for (int k = d->nkeys - 1; k >= 0; k--) {
			P.keys[k].mult = mult;
			if (domain <= 1e18)
				mult *= L.key_range[k];
		}
	}

	// sums
	P.nsums = 0;
	L.sum_of_aggr.assign(d->naggrs, -1);
	for (int a = 0; a < d->naggrs; a++) {
		if (d->aggrs[a].fn != VH_A_SUM)
			continue;
		SumArg& s = P.sums[P.nsums];
		memset(&s, 0, sizeof(s));
		const int root = d->aggrs[a].expr;
		// every node's true value must fit its typed width, as 64-bit
		// evaluation is then exact (no wrap in the reference either)
		for (int n = 0; n <= root; n++) {
			const TypeProps& t = L.ty.nodes[n];
			if (!fits_type(t.type, t.emin, t.emax))
				return fail(VH_ERR_UNSUPPORTED, "aggregate %d: node %d may wrap its type", a, n);
		}
		if (!product_terms(d, root, s))
			return fail(VH_ERR_UNSUPPORTED, "aggregate %d: not a product of linear terms", a);
		const int at = L.ty.nodes[root].type;
		s.arg_signed = type_signed(at);
		s.res128 = L.ty.aggrs[a].type == VH_I128;
		s.wide = s.res128 && (at == VH_U64 || at == VH_I64);
		const TypeProps& rt = L.ty.nodes[root];
		s.narrow = s.arg_signed ? (rt.emin >= -2147483648.0 && rt.emax <= 2147483647.0)
			: (rt.emin >= 0.0 && rt.emax <= 4294967295.0);
		L.sum_of_aggr[a] = P.nsums;
		P.nsums++;
	}

	// partitioned group-by tuple record: the group index (u32) at byte 0,
	// the 4-byte arguments after it, then the 8-byte ones, 8-aligned
	{
		int off = 4;
		for (int s2 = 0; s2 < P.nsums; s2++)
			if (P.sums[s2].narrow) {
				P.sums[s2].toff = off;
				off += 4;
			}
		off = (off + 7) & ~7;
		for (int s2 = 0; s2 < P.nsums; s2++)
			if (!P.sums[s2].narrow) {
				P.sums[s2].toff = off;
				off += 8;
			}
		P.grec = (off + 7) & ~7;
	}

	// group-by mode
	int mode = cfg ? cfg->group_mode : -1;
	const size_t lds_budget = 64 * 1024;
	const double lds_need = domain * P.nsums * 12 + domain * 4 + 16;
	const bool reg_fits = domain <= kRegSlots && P.nsums <= kRegSums &&
		domain * P.nsums * 256 * 12 + domain * 256 * 4 <= 64 * 1024;
	if (mode == VH_GROUP_DENSE)   // dense LDS state: lane-private if it fits, else shared
		mode = reg_fits ? GB_REG : GB_LDS;
	if (mode < 0) {
		// built-in heuristic; the autotuner and the learned selector override
		// it.  A shared LDS table beats lane-private accumulators even at a
		// handful of groups: the latter's LDS footprint (groups x sums x 12
		// bytes per lane) caps the workgroups a CU can hold.  When a group
		// key is the payload of a filtered build, the groups present are
		// usually far fewer than the domain of the payload columns (star:
		// 175 of 4,375), so the LDS hash table (which adds the groups it
		// cannot hold into global memory) is taken up to 16 K; otherwise
		// only while every group fits it.  The shared table is taken up to
		// 64 KiB (two 256-thread workgroups per CU, the default grid): a
		// 1,280-group pipeline of 10M rows spent 0.69 ms partitioned and
		// aggregated by 20 workgroups.
		bool keys_filtered = false;
		for (int k = 0; k < P.nkeys; k++)
			if (P.keys[k].r.kind == ROLE_PAY && d->joins[P.keys[k].r.slot].build.npreds > 0)
				keys_filtered = true;
		const double hash_limit = (keys_filtered ? 16.0 : 1.0) * kLdsHashSlots;
		if (lds_need <= (double)lds_budget)
			mode = GB_LDS;
		else if (domain <= hash_limit && P.nsums <= 4)
			mode = GB_LDSHASH;
		else if (part_fits(domain, P.nsums) && P.rows < ((int64_t)1 << 32))
			mode = GB_PART;
		else if (domain <= (double)(1 << 20) && P.nsums <= 4)
			mode = GB_LDSHASH;
		else if (domain <= (double)(1 << 24))
			mode = GB_GLOBAL;
		else
			mode = GB_HASH;
	}
	if (mode == GB_LDSHASH && !(domain <= (double)(1 << 26) && P.nsums <= 4))
		return fail(VH_ERR_UNSUPPORTED, "LDS hash group-by needs a dense domain <= 2^26 and <= 4 sums");
	if (mode == GB_REG && !(domain <= kRegSlots && P.nsums <= kRegSums &&
		domain * P.nsums * 256 * 12 + domain * 256 * 4 <= 64 * 1024))
		return fail(VH_ERR_UNSUPPORTED, "lane-private group-by needs <= %d groups, <= %d sums and "
			"<= 64 KiB of LDS", kRegSlots, kRegSums);
	if (mode == GB_LDS && lds_need > lds_budget)
		return fail(VH_ERR_UNSUPPORTED, "LDS group-by needs %.0f bytes", lds_need);
	if (mode == GB_GLOBAL && domain > (double)(1 << 26))
		return fail(VH_ERR_UNSUPPORTED, "dense global group-by domain %.0f too large", domain);
	if (mode == GB_HASH && bits > 63)
		return fail(VH_ERR_UNSUPPORTED, "packed group key needs %d bits", bits);
	if (mode == GB_PART && !part_fits(domain, P.nsums))
		return fail(VH_ERR_UNSUPPORTED, "partitioned group-by needs <= %d partitions of the dense domain "
			"and <= 4 sums", kMaxGroupParts);
	// tuple positions of the partitioned group-by (region bases, run offsets,
	// partition totals) are 32-bit
	if (mode == GB_PART && P.rows >= ((int64_t)1 << 32))
		return fail(VH_ERR_UNSUPPORTED, "partitioned group-by over %lld probe rows: at most 2^32 - 1",
			(long long)P.rows);
	if (mode < GB_REG || (mode > GB_LDSHASH && mode != GB_PART))
		return fail(VH_ERR_INVALID, "group mode %d", mode);
	// radix-partitioned probe: the last join's table is built partitioned
	// and the group-by keeps its state per workgroup or in global memory
	// (the join kernel's workgroups are partitions, not row slabs)
	if (cfg && cfg->block != 0 && cfg->block != 256 && cfg->block != 512 && cfg->block != 1024)
		return fail(VH_ERR_INVALID, "block %d: 256, 512 or 1024", cfg->block);
	const int rcfg = cfg ? cfg->radix : 0;
	if (rcfg < -1 || rcfg > 1)
		return fail(VH_ERR_INVALID, "radix %d", rcfg);
	P.radix = 0;
	if (rcfg == 1 && P.njoins > 0 && L.builds[P.njoins - 1].hashed && mode != GB_PART) {
		P.radix = 1;
		if (mode == GB_REG)
			mode = GB_LDS;
	}
	P.mode = mode;
	if (mode == GB_HASH) {
		// groups <= min(domain, tuples); tuples <= probe rows unless a JoinN fans out
		double bound = std::min(domain, (double)std::max<int64_t>(P.rows, 1));
		uint64_t cap = 1024;
		while ((double)cap < 2 * bound && cap < (1ull << 30))
			cap <<= 1;
		P.hmask = cap - 1;
		P.domain = 0;
		L.groups = cap;
	} else {
		P.domain = (uint64_t)domain;
		L.groups = P.domain;
	}
	// exact 128-bit bookkeeping is needed when a sum is i128-typed (global
	// atomics) or its lanes can carry (wide sums)
	L.wide = false;
	for (int s = 0; s < P.nsums; s++)
		L.wide |= P.sums[s].res128 != 0;
	P.wide = L.wide ? 1 : 0;

	// launch shape: one resident wave of workgroups, fewer when there are not
	// 1024 rows per workgroup.  The generated kernel runs best with 2 per CU
	// (measured: c1 probe 35 us at 512 workgroups, 37.6 at 1024, 41 at 2048 on
	// MI355X; each workgroup pays a fixed start-up and partials write), the
	// generic kernel with 4.
	int grid = cfg && cfg->grid > 0 ? cfg->grid : 0;
	bool hash_join = false;
	for (int j = 0; j < P.njoins; j++)
		hash_join = hash_join || !L.builds[j].dir_w;
	if (!grid) {
		int64_t want = (P.rows + 1023) / 1024;
		// lookups in a hash table too large for L2 wait on the Infinity
		// Cache: twice the workgroups hide more of that latency (sparse-key
		// join, 1.5M x 6M: probe 86 us at 2 per CU, 79 at 4, both 4 rows
		// per lane)
		const int per_cu = (cfg && cfg->kernel == VH_KERNEL_GENERIC) || hash_join ? 4 : 2;
		grid = (int)std::max<int64_t>(1, std::min<int64_t>(want, (int64_t)per_cu * device_cus()));
	}
	if (P.mode == GB_PART) {
		// a contiguous slab of rows per workgroup, whole chunks of 1024 rows
		grid = std::min(grid, kMaxGroupRegions);
		const int64_t rows = std::max<int64_t>(P.rows, 1);
		P.gregion_rows = ((rows + grid - 1) / grid + 1023) / 1024 * 1024;
		grid = (int)((rows + P.gregion_rows - 1) / P.gregion_rows);
		P.gstage = part_stage(P);
		part_partitions(P, 256);
	}
	L.grid = grid;
	// dense LDS group-by: workgroups leave partials for merge_kernel, or add
	// into the result with atomics.  The atomics all land on a few words at
	// the end of the launch and cost more than the merge launch unless there
	// are only a few thousand of them.
	const int merge = cfg ? cfg->merge : 0;
	if (merge < 0 || merge > 2)
		return fail(VH_ERR_INVALID, "merge %d", merge);
	if (P.mode == GB_REG || P.mode == GB_LDS)
		P.flush = merge == 2 || (merge == 0 && (double)grid * domain * (P.nsums + 1) <= 16384.0) || P.radix;
	L.build_grid.assign(d->njoins, 0);
	for (int j = 0; j < d->njoins; j++)
		L.build_grid[j] = (int)std::max<int64_t>(1, std::min<int64_t>((d->joins[j].build.rows + 1023) / 1024, 4096));
	return VH_OK;
}

// ------------------------------------------------------------------------
// query: device state, run, result
// ------------------------------------------------------------------------

Query::~Query()
{
	for (void* p : allocs)
		if (p)
			(void)hipFree(p);
}

int Query::alloc(void** p, size_t bytes)
{
	*p = nullptr;
	if (!bytes)
		bytes = 16;
	hipError_t e = hipMalloc(p, bytes);
	if (e != hipSuccess)
		return fail(VH_ERR_NOMEM, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
	allocs.push_back(*p);
	return VH_OK;
}

static int device_lds_per_block()
{
	static int bytes = 0;
	if (!bytes) {
		int dev = 0, n = 0;
		if (hipGetDevice(&dev) == hipSuccess &&
			hipDeviceGetAttribute(&n, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess && n > 0)
			bytes = n;
		else
			bytes = 64 * 1024;
	}
	return bytes;
}

// Join tables the generated probe kernel copies into LDS, in join order
// (the first joins see the most rows): the whole dir[] while it fits what is
// left of the budget, else its key bitmap when the build is filtered or its
// keys are sparse, so that most lookups of keys without a row end in LDS.
// Direct tables only; the generic kernel ignores these fields.
static void assign_lds_tables(Lowered& L, const vh_pipeline_desc* d, size_t budget)
{
	Plan& P = L.plan;
	const size_t kLdsTableBudget = budget;
	size_t used = groupby_lds_bytes(P);
	for (int j = 0; j < P.njoins; j++) {
		JoinT& J = P.joins[j];
		J.lds = LDS_NONE;
		J.lds_off = 0;
		J.bits = nullptr;
		if (!J.dir_w)
			continue;
		const double keys = (double)(J.kmax - J.kmin) + 1;
		const bool misses = d->joins[j].build.npreds > 0 || keys > 2.0 * (double)d->joins[j].build.rows;
		J.lds = LDS_DIR;
		if (used + join_lds_bytes(J) > kLdsTableBudget) {
			J.lds = LDS_BITS;
			if (!misses || used + join_lds_bytes(J) > kLdsTableBudget) {
				J.lds = LDS_NONE;
				continue;
			}
		}
		J.lds_off = (u32)used;
		used += join_lds_bytes(J);
	}
}

// With join tables in LDS the generated kernel runs 1024-thread workgroups,
// one per CU: sixteen waves share one copy of the tables (with 256 threads,
// 60 KB of tables left two workgroups, two waves per SIMD, for a chain of
// dependent lookups), and the tables may use the whole LDS of a workgroup
// (160 KB on gfx950).  Lane-private and partitioned group-bys keep 256
// threads: their LDS and row slabs are laid out for that.
// Radix-partitioned probe: workgroups of 512 threads (or the configured
// size) take chunks of U * block rows whose tuples fit 64 KiB of LDS
// staging, two workgroups per CU (U = 8 for tuples of 12 bytes, a key
// offset and an int64).  A probe partition is the largest aligned block of
// the table, of at least one of its build's sub-tables and at most 2^14
// slots (128 KiB), whose LDS copy fits beside the group-by state and the
// run tables; the join kernel runs one 1024-thread workgroup per partition.
// Where no such layout exists the one-pass probe runs instead (P.radix = 0).
static void choose_radix_shape(Lowered& L, const vh_config* cfg)
{
	Plan& P = L.plan;
	const size_t lds_max = 160 * 1024;   // gfx950: LDS per workgroup
	const size_t stage_max = 64 * 1024;
	auto r16 = [](size_t x) { return (x + 15) & ~(size_t)15; };
	const JoinT& J = P.joins[P.njoins - 1];
	std::vector<int> widths;
	const int tb = jit_radix_values(P, widths);
	const int block = (cfg && cfg->block > 0) ? cfg->block : 512;
	int U = jit_unroll(cfg);
	if (!U) {
		// a tuple's bytes in LDS: key offset, values and a 2-byte sort position
		const size_t per_row = (size_t)tb + 2;
		U = per_row * 8 * block <= stage_max ? 8 : (per_row * 4 * block <= stage_max ? 4 : 2);
	}
	const u64 CR = (u64)U * (u64)block;
	const u64 RC = ((u64)std::max<int64_t>(P.rows, 0) + CR - 1) / CR + 1;
	const int tshift = 63 - __builtin_clzll(J.mask + 1);
	const int bshift = 63 - __builtin_clzll(J.smask + 1);
	const size_t gb = r16(groupby_lds_bytes(P));
	int rps = -1;
	for (int sh = std::min(tshift, 14); sh >= bshift; sh--)
		if (gb + ((size_t)8 << sh) + (2 * RC + 2) * 4 <= lds_max) {
			rps = sh;
			break;
		}
	bool ok = rps >= 0 && widths.size() <= (size_t)kRadixMaxVals && RC <= (u64)kRadixMaxChunks &&
		RC * CR < ((u64)1 << 32);
	if (ok) {
		P.rparts = (int)((J.mask + 1) >> rps);
		P.rlds_a = (u32)r16(pipeline_lds_bytes(P));
		const size_t stage = jit_radix_stage_bytes(P, CR);
		ok = P.rparts <= 4096 && P.rlds_a + stage <= lds_max;
		if (ok) {
			P.rpshift = rps;
			P.rchunk = (i64)CR;
			P.rchunks = (int)RC;
			P.rw = (int)widths.size();
			P.rlds_b = (u32)gb;
			L.radix_lds_a = P.rlds_a + stage;
			L.radix_lds_b = gb + ((size_t)8 << rps) + (2 * RC + 2) * 4;
			L.radix_unroll = U;
			L.radix_block = 1024;
			L.radix_grid = P.rparts;
			L.block = block;
			if (!(cfg && cfg->grid > 0)) {
				// workgroups a CU holds: its LDS, and 32 waves
				const int per_cu = (int)std::max<size_t>(1, std::min<size_t>(lds_max / L.radix_lds_a,
					(size_t)(2048 / block)));
				L.grid = (int)std::max<u64>(1, std::min<u64>(RC, (u64)per_cu * (u64)device_cus()));
			}
		}
	}
	if (!ok)
		P.radix = 0;
}

void choose_probe_shape(Lowered& L, const vh_pipeline_desc* d, const vh_config* cfg)
{
	Plan& P = L.plan;
	const int lowered_grid = L.grid;   // the per-workgroup partials are sized for it
	const bool fixed = P.mode == GB_REG || P.mode == GB_PART;
	const size_t budget = fixed ? 64 * 1024 : std::min<size_t>((size_t)device_lds_per_block(), 160 * 1024);
	assign_lds_tables(L, d, budget);
	L.block = 256;
	bool tables = false;
	for (int j = 0; j < P.njoins; j++)
		tables = tables || P.joins[j].lds != LDS_NONE;
	if (tables && !fixed) {
		L.block = 1024;
		if (!(cfg && cfg->grid > 0))
			L.grid = std::max(1, std::min(L.grid, device_cus()));
	}
	// an explicit workgroup size (waves per CU = block / 64 x workgroups per
	// CU); the automatic grid then keeps the same threads per CU, at least
	// one workgroup per CU
	const int want = cfg ? cfg->block : 0;
	if (want > 0 && !fixed && want != L.block) {
		const int per_cu_before = L.block;
		L.block = want;
		if (!(cfg && cfg->grid > 0)) {
			const int64_t threads = (int64_t)L.grid * per_cu_before;
			L.grid = (int)std::max<int64_t>(1, std::min<int64_t>(threads / want, 4LL * device_cus()));
			L.grid = std::max(L.grid, std::min(device_cus(), (int)((P.rows + 1023) / 1024)));
			L.grid = std::min(L.grid, lowered_grid);
		}
	}
	if (P.radix)
		choose_radix_shape(L, cfg);
}

// From the generated probe kernel(s) to the generic one, which computes the
// same result: it reads no LDS join tables, runs 256-thread workgroups over
// the grid the lowering chose, and has no radix-partitioned probe.
static int switch_to_generic(Query* q)
{
	Lowered& L = q->L;
	Plan& P = L.plan;
	q->jit_fn = nullptr;
	q->jit_src.clear();
	q->jit_join_fn = nullptr;
	q->jit_join_src.clear();
	P.radix = 0;
	for (int j = 0; j < P.njoins; j++) {
		P.joins[j].lds = LDS_NONE;
		P.joins[j].lds_off = 0;
		P.joins[j].bits = nullptr;
		L.builds[j].bits = nullptr;
	}
	L.block = 256;
	L.grid = q->generic_grid;
	HIP_TRY(hipMemcpy(q->dplan, &P, sizeof(Plan), hipMemcpyHostToDevice));
	if (P.njoins)
		HIP_TRY(hipMemcpy(q->dbuilds, L.builds.data(), P.njoins * sizeof(BuildPlan), hipMemcpyHostToDevice));
	return VH_OK;
}

int query_create(Query* q, bool single)
{
	Lowered& L = q->L;
	Plan& P = L.plan;
	int rc;
	// flags (kFlagWords u32): [0] error bits, [2 + 2j] repeated keys in
	// join j's build, each in a 64-bit word of its own so the clear pass can
	// zero them separately
	if ((rc = q->alloc((void**)&q->flags, kFlagWords * sizeof(u32))))
		return rc;
	HIP_TRY(hipMemset(q->flags, 0, kFlagWords * sizeof(u32)));
	P.error = q->flags;
	for (int j = 0; j < P.njoins; j++) {
		if ((rc = q->alloc((void**)&q->slots[j], L.capacity[j] * sizeof(u64))))
			return rc;
		// the run's clear pass refills the slots only after a run that used
		// them, so they start out empty here
		HIP_TRY(hipMemset(q->slots[j], 0xFF, L.capacity[j] * sizeof(u64)));
		P.joins[j].slots = q->slots[j];
		P.joins[j].dups = dups_flag(q, j);
		if (L.dir_words[j]) {
			if ((rc = q->alloc((void**)&q->dir[j], L.dir_words[j] * sizeof(u64))))
				return rc;
			HIP_TRY(hipMemset(q->dir[j], 0xFF, L.dir_words[j] * sizeof(u64)));
			P.joins[j].dir = q->dir[j];
			L.builds[j].dir = q->dir[j];
		}
		BuildPlan& B = L.builds[j];
		if (B.hashed) {
			if ((rc = q->alloc((void**)&q->ovf[j], L.capacity[j] * sizeof(u64))))
				return rc;
			HIP_TRY(hipMemset(q->ovf[j], 0xFF, L.capacity[j] * sizeof(u64)));
			B.ovf = q->ovf[j];
			B.spill = spill_flag(q, j);
			P.joins[j].ovf = q->ovf[j];
			P.joins[j].spill = spill_flag(q, j);
		}
		if (B.nparts) {
			const size_t g = (size_t)std::max(B.nchunks, 1);
			if ((rc = q->alloc((void**)&B.pbuf, g * kBuildChunkRows * sizeof(u64))) ||
				(rc = q->alloc((void**)&B.pcnt, g * (size_t)B.nparts * sizeof(u32))) ||
				(rc = q->alloc((void**)&B.poff, g * (size_t)B.nparts * sizeof(u32))))
				return rc;
		}
		L.builds[j].slots = q->slots[j];
		L.builds[j].dups = dups_flag(q, j);
		L.builds[j].error = q->flags;
	}
	const u64 groups = L.groups;
	const int nsums = std::max(P.nsums, 1);
	if ((rc = q->alloc((void**)&P.out_lo, groups * nsums * sizeof(u64))))
		return rc;
	if ((rc = q->alloc((void**)&P.out_hi, groups * nsums * sizeof(i64))))
		return rc;
	if ((rc = q->alloc((void**)&P.out_cnt, groups * sizeof(u64))))
		return rc;
	if (P.mode == GB_REG || P.mode == GB_LDS) {
		const u64 W = 2 * P.domain * P.nsums + P.domain;
		if ((rc = q->alloc((void**)&P.partials, (u64)L.grid * W * sizeof(u64))))
			return rc;
	}
	if (P.mode == GB_PART) {
		const u64 rows = (u64)std::max<int64_t>(P.rows, 1);
		const u64 regions = (rows + (u64)P.gregion_rows - 1) / (u64)P.gregion_rows;
		if ((rc = q->alloc((void**)&P.gtup0, rows * (u64)P.grec)) ||
			(rc = q->alloc((void**)&P.gtup1, rows * (u64)P.grec)) ||
			(rc = q->alloc((void**)&P.gend, regions * (u64)P.gparts * sizeof(u32))))
			return rc;
	}
	if (P.mode == GB_HASH) {
		if ((rc = q->alloc((void**)&P.hkeys, groups * sizeof(u64))))
			return rc;
	}
	// the clear pass of every run, one launch: join tables, atomically
	// accumulated outputs, error bits.  A partitioned build rewrites its
	// whole direct table and clears its own repeated-key flag
	// (part_scatter_kernel), so only its overflow slots are cleared here, and
	// only after a run that used them; other joins clear their flag here.
	// two or more partitioned builds: their chunks and partitions become the
	// work items of one scatter and one slice launch
	int npart = 0;
	for (int j = 0; j < P.njoins; j++)
		npart += L.builds[j].nparts > 0;
	if (single && npart >= 2) {
		if ((rc = q->alloc((void**)&q->dctl, 2 * sizeof(u32))))
			return rc;
		std::vector<BatchItem> items;
		for (int j = 0; j < P.njoins; j++)
			for (int c = 0; c < L.builds[j].nchunks && L.builds[j].nparts; c++)
				items.push_back({0, 0, j, 0, c, 0});
		q->nscatter = (u32)items.size();
		for (int j = 0; j < P.njoins; j++) {
			for (int p = 0; p < L.builds[j].nparts; p++)
				items.push_back({0, 0, j, 0, p, 0});
			if (L.builds[j].nparts)
				q->slice_lds = std::max(q->slice_lds, part_slice_lds_bytes(L.builds[j]));
		}
		q->nslice = (u32)items.size() - q->nscatter;
		if ((rc = q->alloc((void**)&q->dparts, items.size() * sizeof(BatchItem))))
			return rc;
		HIP_TRY(hipMemcpy(q->dparts, items.data(), items.size() * sizeof(BatchItem), hipMemcpyHostToDevice));
	}
	std::vector<FillSeg> segs;
	segs.push_back({(u64*)q->flags, 1, 0, nullptr});
	if (q->dctl)
		segs.push_back({(u64*)q->dctl, 1, 0, nullptr});
	for (int j = 0; j < P.njoins; j++) {
		if (L.builds[j].hashed) {
			// the slices store every slot; only the overflow table is
			// refilled, after a run that spilled into it
			segs.push_back({q->ovf[j], L.capacity[j], ~0ull, spill_flag(q, j)});
		} else if (L.builds[j].nparts) {
			segs.push_back({q->slots[j], L.capacity[j], ~0ull, dups_flag(q, j)});
		} else {
			segs.push_back({(u64*)dups_flag(q, j), 1, 0, nullptr});
			segs.push_back({q->slots[j], L.capacity[j], ~0ull, nullptr});
			if (q->dir[j])
				segs.push_back({q->dir[j], L.dir_words[j], ~0ull, nullptr});
		}
	}
	if (P.mode == GB_GLOBAL || P.mode == GB_HASH || P.mode == GB_LDSHASH || P.mode == GB_PART || P.flush) {
		segs.push_back({P.out_lo, groups * nsums, 0, nullptr});
		segs.push_back({(u64*)P.out_hi, groups * nsums, 0, nullptr});
		segs.push_back({P.out_cnt, groups, 0, nullptr});
		if (P.mode == GB_HASH)
			segs.push_back({P.hkeys, groups, ~0ull, nullptr});
	}
	q->nclear = (int)segs.size();
	if ((rc = q->alloc((void**)&q->dclear, segs.size() * sizeof(FillSeg))))
		return rc;
	HIP_TRY(hipMemcpy(q->dclear, segs.data(), segs.size() * sizeof(FillSeg), hipMemcpyHostToDevice));

	// join tables the generated probe kernel keeps in LDS, its workgroup size
	q->generic_grid = L.grid;
	if (single && q->cfg.kernel != VH_KERNEL_GENERIC) {
		choose_probe_shape(L, &q->desc, &q->cfg);
		for (int j = 0; j < P.njoins; j++) {
			if (P.joins[j].lds != LDS_BITS)
				continue;
			if ((rc = q->alloc((void**)&q->bits[j], join_lds_bytes(P.joins[j]))))
				return rc;
			P.joins[j].bits = q->bits[j];
			if (L.builds[j].nparts)
				L.builds[j].bits = q->bits[j];   // written with the slices
		}
	}

	// radix-partitioned probe: generated kernels only (its shape is set by
	// choose_probe_shape); its tuple arrays and run table
	if (P.radix && !(single && q->cfg.kernel != VH_KERNEL_GENERIC))
		P.radix = 0;
	if (P.radix) {
		std::vector<int> widths;
		jit_radix_values(P, widths);
		const u64 tuples = (u64)P.rchunks * (u64)P.rchunk;
		if ((rc = q->alloc((void**)&P.rkey, tuples * sizeof(u32))) ||
			(rc = q->alloc((void**)&P.rend, (u64)P.rchunks * (u64)P.rparts * sizeof(u32))))
			return rc;
		for (size_t k = 0; k < widths.size(); k++)
			if ((rc = q->alloc(&P.rval[k], tuples * (u64)widths[k])))
				return rc;
	}

	// device copies of the plans
	if ((rc = q->alloc((void**)&q->dplan, sizeof(Plan))))
		return rc;
	HIP_TRY(hipMemcpy(q->dplan, &P, sizeof(Plan), hipMemcpyHostToDevice));
	if (P.njoins) {
		if ((rc = q->alloc((void**)&q->dbuilds, P.njoins * sizeof(BuildPlan))))
			return rc;
		HIP_TRY(hipMemcpy(q->dbuilds, L.builds.data(), P.njoins * sizeof(BuildPlan), hipMemcpyHostToDevice));
		for (int j = 0; j < P.njoins; j++)
			q->dbuild[j] = q->dbuilds + j;
	}
	// the generated probe kernel, unless the caller asked for the generic one
	const int kernel = q->cfg.kernel;
	if (kernel < VH_KERNEL_AUTO || kernel > VH_KERNEL_JIT_PIPE)
		return fail(VH_ERR_INVALID, "kernel %d", kernel);
	if (single && kernel != VH_KERNEL_GENERIC) {
		q->jit_src = jit_source(P, jit_eager(kernel), P.radix ? L.radix_unroll : jit_unroll(&q->cfg),
			jit_pipeline(kernel), L.block);
		std::string log;
		rc = jit_compile(q->jit_src, &q->jit_fn, log);
		if (!rc && P.radix) {
			q->jit_join_src = jit_join_source(P, jit_unroll(&q->cfg), L.radix_block);
			rc = jit_compile(q->jit_join_src, &q->jit_join_fn, log, "vh_generated_join");
			if (rc && kernel == VH_KERNEL_AUTO) {
				// the one-pass probe computes the same result
				P.radix = 0;
				q->jit_join_fn = nullptr;
				q->jit_join_src.clear();
				q->jit_src = jit_source(P, jit_eager(kernel), jit_unroll(&q->cfg), jit_pipeline(kernel), L.block);
				rc = jit_compile(q->jit_src, &q->jit_fn, log);
				if (!rc)
					HIP_TRY(hipMemcpy(q->dplan, &P, sizeof(Plan), hipMemcpyHostToDevice));
			}
		}
		if (rc && kernel != VH_KERNEL_AUTO)
			return rc;
		if (rc)   // AUTO: the generic kernel computes the same result
			return switch_to_generic(q);
	}
	return VH_OK;
}

int query_run(Query* q, hipStream_t s, int phase)
{
	Lowered& L = q->L;
	Plan& P = L.plan;
	if (phase < -1 || phase > 3)
		return fail(VH_ERR_INVALID, "phase %d", phase);
	if (phase == -1 || phase == 0)
		HIP_TRY(launch_multi_fill(q->dclear, q->nclear, s));
	if (phase == -1 || phase == 1) {
		if (q->dparts) {
			// a persistent workgroup per item slot: 48 KiB of LDS, three per CU
			const int cap = 3 * device_cus();
			HIP_TRY(launch_batch_part(false, q->dbuilds, q->dparts, 0, q->nscatter, q->dctl,
				(int)std::max<u32>(1, std::min<u32>(q->nscatter, (u32)cap)), 0, s));
			HIP_TRY(launch_batch_part(true, q->dbuilds, q->dparts, q->nscatter, q->nslice, q->dctl + 1,
				(int)std::max<u32>(1, std::min<u32>(q->nslice, (u32)cap)), q->slice_lds, s));
		}
		for (int j = 0; j < P.njoins; j++) {
			if (L.builds[j].nparts && !q->dparts)
				HIP_TRY(launch_part_build(L.builds[j], q->dbuild[j], s));
			else if (!L.builds[j].nparts && L.builds[j].rows > 0)
				HIP_TRY(launch_build(q->dbuild[j], L.build_grid[j], s));
		}
		for (int j = 0; j < P.njoins; j++)
			if (P.joins[j].lds == LDS_BITS && !L.builds[j].nparts)
				HIP_TRY(launch_dir_bits(q->dir[j], P.joins[j].dir_w, (u64)(P.joins[j].kmax - P.joins[j].kmin) + 1,
					q->bits[j], s));
	}
	if (phase == -1 || phase == 2) {
		if (q->jit_join_fn) {
			// radix-partitioned probe: partitioned tuples, then one join
			// workgroup per partition (the stamps span both launches)
			HIP_TRY(launch_generated(q->jit_fn, P, q->dplan, L.grid, L.block, s, q->probe_start, nullptr,
				(long)L.radix_lds_a));
			HIP_TRY(launch_generated(q->jit_join_fn, P, q->dplan, L.radix_grid, L.radix_block, s, nullptr,
				q->probe_stop, (long)L.radix_lds_b));
		} else if (q->jit_fn)
			HIP_TRY(launch_generated(q->jit_fn, P, q->dplan, L.grid, L.block, s, q->probe_start, q->probe_stop));
		else
			HIP_TRY(launch_pipeline(P, q->dplan, L.grid, s));
	}
	if (phase == -1 || phase == 3) {
		HIP_TRY(launch_merge(P, q->dplan, L.grid, s));
		q->ran = true;
	}
	return VH_OK;
}

// voila_cast<T>::to_cstr (runtime.hpp:348-387): magnitude digits, then '-'.
static void append_i128(std::string& out, __int128 v)
{
	char buf[41];
	unsigned __int128 m = v < 0 ? (unsigned __int128)(-v) : (unsigned __int128)v;
	char* d = buf + sizeof(buf);
	do {
		*--d = (char)('0' + (int)(m % 10));
		m /= 10;
	} while (m);
	if (v < 0)
		*--d = '-';
	out.append(d, (size_t)(buf + sizeof(buf) - d));
}

static __int128 typed_value(int t, u64 lo, i64 hi)
{
	switch (t) {
	case VH_U64: return (__int128)lo;
	case VH_I64: return (__int128)(i64)lo;
	case VH_I128: return (__int128)(((unsigned __int128)(u64)hi << 64) | lo);
	default: return type_signed(t) ? (__int128)(i64)lo : (__int128)lo;
	}
}

int query_result(Query* q, hipStream_t s, std::string& text, int64_t& nrows)
{
	if (!q->ran)
		return fail(VH_ERR_STATE, "query has not run");
	Lowered& L = q->L;
	Plan& P = L.plan;
	const vh_pipeline_desc& d = q->desc;
	HIP_TRY(hipStreamSynchronize(s));
	u32 flags[kFlagWords];
	HIP_TRY(hipMemcpy(flags, q->flags, sizeof(flags), hipMemcpyDeviceToHost));
	if (flags[0] & 4u)
		return fail(VH_ERR_INVALID, "a build key lies outside its column's [min, max]");
	if (flags[0] & 1u)
		return fail(VH_ERR_CAPACITY, "group hash table full");
	if (flags[0] & 8u)
		return fail(VH_ERR_INVALID, "a group key lies outside its column's [min, max]");
	if (flags[0] & 32u)
		return fail(VH_ERR_CAPACITY, "a hash table probe sequence exceeded its bound (damaged or full table)");
	if (flags[0] & 16u) {
		// a JoinN before the last join matched a repeated build key: the
		// generated kernels take first matches only there, the generic
		// kernel fans out at any join.  Run again with it (and keep it).
		if (!q->jit_fn || q->fanout_rerun)
			return fail(VH_ERR_UNSUPPORTED, "a JoinN before the last join matched a repeated build key: batched "
				"generated kernels take first matches there (give the pipeline the generic kernel)");
		int rc = switch_to_generic(q);
		if (rc)
			return rc;
		q->fanout_rerun = true;
		if ((rc = query_run(q, s, -1)))
			return rc;
		return query_result(q, s, text, nrows);
	}
	const u64 groups = L.groups;
	const int nsums = std::max(P.nsums, 1);
	std::vector<u64> lo(groups * nsums), cnt(groups), hk;
	std::vector<i64> hi(groups * nsums);
	HIP_TRY(hipMemcpy(lo.data(), P.out_lo, lo.size() * 8, hipMemcpyDeviceToHost));
	HIP_TRY(hipMemcpy(hi.data(), P.out_hi, hi.size() * 8, hipMemcpyDeviceToHost));
	HIP_TRY(hipMemcpy(cnt.data(), P.out_cnt, cnt.size() * 8, hipMemcpyDeviceToHost));
	if (P.mode == GB_HASH) {
		hk.resize(groups);
		HIP_TRY(hipMemcpy(hk.data(), P.hkeys, hk.size() * 8, hipMemcpyDeviceToHost));
	}
	text.clear();
	nrows = 0;
	for (u64 g = 0; g < groups; g++) {
		if (!cnt[g])
			continue;   // a group exists iff its count > 0 (relalg_translator.cpp:698-701)
		bool first = true;
		for (int k = 0; k < d.nkeys; k++) {
			int64_t v;
			if (P.mode == GB_HASH) {
				const u64 m = L.key_bits[k] >= 64 ? ~0ull : ((1ull << L.key_bits[k]) - 1);
				v = P.keys[k].min + (int64_t)((hk[g] >> P.keys[k].shift) & m);
			} else {
				v = P.keys[k].min + (int64_t)((g / P.keys[k].mult) % L.key_range[k]);
			}
			if (!first)
				text.push_back('|');
			first = false;
			append_i128(text, (__int128)v);
		}
		for (int a = 0; a < d.naggrs; a++) {
			__int128 v;
			if (d.aggrs[a].fn == VH_A_COUNT) {
				v = (__int128)cnt[g];
			} else {
				const int s2 = L.sum_of_aggr[a];
				v = typed_value(L.ty.aggrs[a].type, lo[s2 * groups + g], hi[s2 * groups + g]);
			}
			if (!first)
				text.push_back('|');
			first = false;
			append_i128(text, v);
		}
		text.push_back('\n');
		nrows++;
	}
	return VH_OK;
}

} // namespace vh
