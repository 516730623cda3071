// Generated probe kernels.
//
// The reference turns every query into C++ for its flavor and compiles it
// with g++ (codegen.cpp:211-395, compiler.cpp:19-55).  Here a lowered Plan
// becomes HIP source for one fused kernel: the column types, the predicate,
// join and aggregate sequence, each join's table layout and the group-by
// strategy are fixed in the text, while pointers and constants stay in the
// Plan the kernel reads (so queries that differ only in data share one
// compiled kernel).  hipRTC compiles it for the device's gfx target; the
// code object is cached per (target, source) for the life of the process.
//
// The generated code calls the same device helpers as the generic kernel
// (vh_fused.hpp: compare, voila_hash, locate / add_count / add_sum,
// lds_layout / lds_init, block_epilogue), with the plan's choices as
// literals, so both compute identical results.  What the generated kernel
// gains: every probe column is loaded for all U rows of a chunk in one go,
// before the predicates, so a chunk costs one round trip to memory for its
// columns plus one per join, and register use is what this plan needs.
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>
#include <hip/hiprtc.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/voila_hip.h"
#include "vh_engine.hpp"
#include "vh_jit.hpp"
#include "vh_kernels.hpp"

namespace vh {

#include "vh_jit_headers.inc"   // kJitHeaderNames / kJitHeaderText (build.py)

namespace {

const char* ctype_of(int t)
{
	switch (t) {
	case VH_U8: return "uint8_t";
	case VH_I8: return "int8_t";
	case VH_U16: return "uint16_t";
	case VH_I16: return "int16_t";
	case VH_U32: return "uint32_t";
	case VH_I32: return "int32_t";
	default: return "i64";   // 64-bit columns are read as i64, like load1
	}
}

int host_width(int t)
{
	switch (t) {
	case VH_U8: case VH_I8: return 1;
	case VH_U16: case VH_I16: return 2;
	case VH_U32: case VH_I32: return 4;
	default: return 8;
	}
}

struct Col {
	const void* ptr;
	int type;
	std::string path;   // expression for the column's address in the kernel
	bool pred;          // read by a predicate (loaded first)
	int stage;          // first use: 0 predicates, j + 1 join j's key, njoins + 1 aggregation
	bool late = false;  // vector loop: loaded for live rows only, before stage `stage`
};

struct Gen {
	const Plan& P;
	bool eager;         // scalar path: load every column before the predicates
	bool pipe;          // vector path: the next chunk's columns are requested
	                    // behind this chunk's last table lookup
	int U;
	bool vec = false;   // vector main loop (decided in source())
	int late_stage = 0; // vector loop: stage before which late columns are loaded (0: none)
	int block = 256;    // threads per workgroup
	bool batch = false; // entry point of a batched launch (vh_generated_batch)
	bool racc = false;  // a dense LDS group-by of one group: every lane sums in
	                    // registers and adds into the LDS table once, at the
	                    // end (64 lanes adding into one LDS word serialise)
	bool radix = false; // P.radix: the probe kernel ends at the last join with
	                    // partitioned tuples (vh_generated), the join kernel
	                    // (vh_generated_join) takes them from there
	std::vector<Col> cols;
	// plan fields the kernel reads, copied into locals before the row loop
	// (the loop's atomics would otherwise make the compiler reload them)
	std::map<std::string, std::string> uni;
	std::string decls;
	std::string out;

	Gen(const Plan& p, bool e, bool pp, int u) : P(p), eager(e), pipe(pp), U(u) {}

	void line(const std::string& s) { out += s; out += '\n'; }

	// Read-only pointer into global memory.  The plan stores generic
	// pointers; loads through them would be flat loads, which also count
	// against the LDS counter and so can only be waited for all together.
	static std::string gptr(const std::string& t)
	{
		return "const __attribute__((address_space(1))) " + t + "*";
	}

	std::string uniform(const std::string& type, const std::string& expr)
	{
		auto it = uni.find(expr);
		if (it != uni.end())
			return it->second;
		const std::string name = "k" + std::to_string(uni.size());
		uni[expr] = name;
		decls += "\t" + type + " const " + name + " = " + expr + ";\n";
		return name;
	}

	// Values a radix probe tuple carries after its key offset: every probe
	// column and earlier join payload the group keys and sum arguments read
	// (the last join's payload comes from the join kernel's lookup).  The
	// tuple's word k + 1 holds value k.
	std::vector<std::pair<Role, std::string>> radix_values() const
	{
		std::vector<std::pair<Role, std::string>> v;
		const int last = P.njoins - 1;
		auto add = [&](const Role& r, const std::string& path) {
			if (r.kind == ROLE_PAY && r.slot == last)
				return;
			if (r.kind != ROLE_COL && r.kind != ROLE_PAY)
				return;
			for (auto& x : v)
				if (x.first.kind == r.kind && (r.kind == ROLE_PAY ? x.first.slot == r.slot
					: (x.first.ptr == r.ptr && x.first.type == r.type)))
					return;
			v.push_back({r, path});
		};
		char b[96];
		for (int k = 0; k < P.nkeys; k++) {
			snprintf(b, sizeof b, "P.keys[%d].r", k);
			add(P.keys[k].r, b);
		}
		for (int si = 0; si < P.nsums; si++)
			for (int t = 0; t < P.sums[si].nterms; t++) {
				const Term& T = P.sums[si].t[t];
				snprintf(b, sizeof b, "P.sums[%d].t[%d]", si, t);
				if (T.sa)
					add(T.a, std::string(b) + ".a");
				if (T.sb)
					add(T.b, std::string(b) + ".b");
			}
		return v;
	}

	// C type and width of a radix tuple value: a probe column's own type,
	// or an earlier join's packed payload word
	static std::string radix_vtype(const Role& r)
	{
		return r.kind == ROLE_PAY ? "u32" : ctype_of(r.type);
	}
	static int radix_vwidth(const Role& r)
	{
		return r.kind == ROLE_PAY ? 4 : host_width(r.type);
	}

	// The probe kernel's LDS staging of a chunk's tuples, from P.rlds_a:
	// counts and cursors per partition, two words (staged count, scan
	// total), then the key offsets and each value as arrays of rchunk
	// entries, every area rounded up to 16 bytes
	void radix_stage_decls(std::string& k) const
	{
		char b[512];
		k += "\tconst int NP = P.rparts;\n\tconst u32 CRr = (u32)P.rchunk;\n";
		k += "\tu32* const rc = (u32*)((char*)lds + P.rlds_a);\n";
		k += "\tu32* const rcur = rc + NP;\n\tu32* const rmeta = rcur + NP;\n";
		k += "\tu64 ro_ = ((u64)P.rlds_a + (2 * (u64)NP + 4) * 4 + 15) & ~15ull;\n";
		k += "\tu32* const skey = (u32*)((char*)lds + ro_);\n\tro_ += ((u64)CRr * 4 + 15) & ~15ull;\n";
		const auto vals = radix_values();
		for (size_t i = 0; i < vals.size(); i++) {
			const std::string t = radix_vtype(vals[i].first);
			snprintf(b, sizeof b, "\t%s* const sv%zu = (%s*)((char*)lds + ro_);\n\tro_ += ((u64)CRr * %d + 15) & ~15ull;\n",
				t.c_str(), i, t.c_str(), radix_vwidth(vals[i].first));
			k += b;
		}
		k += "\tuint16_t* const sperm = (uint16_t*)((char*)lds + ro_);\n\tro_ += ((u64)CRr * 2 + 15) & ~15ull;\n";
		k += "\t(void)ro_;\n";
		k += "\tfor (int i = threadIdx.x; i < NP; i += blockDim.x) rc[i] = 0;\n";
		k += "\tif (threadIdx.x == 0) rmeta[0] = 0;\n";
	}

	// Bytes of that staging for chunks of `chunk` rows
	size_t radix_stage_bytes(u64 chunk) const
	{
		auto r16 = [](u64 x) { return (x + 15) & ~15ull; };
		u64 n = r16((2 * (u64)P.rparts + 4) * 4) + r16(chunk * 4);
		for (const auto& v : radix_values())
			n += r16(chunk * (u64)radix_vwidth(v.first));
		n += r16(chunk * 2);   // the sort's permutation
		return (size_t)n;
	}

	// The flush of chunk `c` (an expression): the staged tuples go to their
	// partitions' runs of the chunk, sorted by partition
	void radix_flush(const std::string& c)
	{
		char b[512];
		const int last = P.njoins - 1;
		snprintf(b, sizeof b, "P.joins[%d].mask", last);
		const std::string mask = uniform("u64", b);
		const std::string rps = uniform("int", "P.rpshift");
		line("\tradix_flush_begin(P, (u32)(" + c + "), rc, rcur, &rmeta[1]);");
		// the staged tuples' sorted positions (an LDS permutation), then
		// coalesced stores of the sorted chunk
		line("\t{");
		line("\t\tconst u32 n_ = rmeta[0];");
		line("\t\tfor (u32 i = threadIdx.x; i < n_; i += blockDim.x)");
		line("\t\t\tsperm[atomicAdd(&rcur[(u32)((voila_hash((u64)skey[i]) & " + mask + ") >> " + rps + ")], 1u)] = (uint16_t)i;");
		line("\t\t__syncthreads();");
		line("\t\tconst u64 cb_ = (u64)(" + c + ") * (u64)CRr;");
		line("\t\tfor (u32 t = threadIdx.x; t < n_; t += blockDim.x) {");
		line("\t\t\tconst u32 i = sperm[t];");
		line("\t\t\t((VH_GLOBAL u32*)P.rkey)[cb_ + t] = skey[i];");
		const auto vals = radix_values();
		for (size_t i = 0; i < vals.size(); i++) {
			snprintf(b, sizeof b, "\t\t\t((VH_GLOBAL %s*)P.rval[%zu])[cb_ + t] = sv%zu[i];", radix_vtype(vals[i].first).c_str(), i, i);
			line(b);
		}
		line("\t\t}");
		line("\t}");
		line("\tradix_flush_end(P, rc, &rmeta[0]);");
	}

	int col_index(const Role& r, const std::string& path, bool pred, int stage = 1 << 20)
	{
		for (size_t i = 0; i < cols.size(); i++)
			if (cols[i].ptr == r.ptr && cols[i].type == r.type) {
				cols[i].pred = cols[i].pred || pred;
				cols[i].stage = std::min(cols[i].stage, stage);
				return (int)i;
			}
		cols.push_back({r.ptr, r.type, path + ".ptr", pred, stage});
		return (int)cols.size() - 1;
	}

	void note_role(const Role& r, const std::string& path, bool pred, int stage)
	{
		if (r.kind == ROLE_COL)
			col_index(r, path, pred, stage);
	}

	// Value of a role for row u as an i64 expression.  tail_pay names the
	// payload variable that replaces join `tail_join`'s in the JoinN tail.
	std::string role(const Role& r, const std::string& path, int tail_join = -1,
		const char* tail_pay = nullptr)
	{
		char b[256];
		switch (r.kind) {
		case ROLE_COL:
			snprintf(b, sizeof b, "((i64)c%d[u])", col_index(r, path, false));
			return b;
		case ROLE_PAY: {
			// payload `item` of the join: the whole entry word for a join
			// with one payload column, else its bit field
			std::string word;
			if (r.slot == tail_join)
				word = tail_pay;
			else {
				snprintf(b, sizeof b, "pay%d[u]", r.slot);
				word = b;
			}
			if (P.joins[r.slot].npay <= 1) {
				snprintf(b, sizeof b, "P.joins[%d].pmin", r.slot);
				return "((i64)" + word + " + " + uniform("i64", b) + ")";
			}
			snprintf(b, sizeof b, "P.joins[%d].pminc[%d]", r.slot, r.item);
			const std::string pmin = uniform("i64", b);
			snprintf(b, sizeof b, "P.joins[%d].pshiftc[%d]", r.slot, r.item);
			const std::string sh = uniform("int", b);
			snprintf(b, sizeof b, "P.joins[%d].pmaskc[%d]", r.slot, r.item);
			const std::string m = uniform("u32", b);
			return "((i64)((" + word + " >> " + sh + ") & " + m + ") + " + pmin + ")";
		}
		case ROLE_CONST:
			return "(" + uniform("i64", path + ".c") + ")";
		default:
			return "((i64)0)";
		}
	}

	// Argument of sum si for row u: the product of its linear terms (u64
	// wrap-around, as eval_sum_u computes it).
	std::string sum_expr(int si, int tail_join, const char* tail_pay)
	{
		char b[256];
		const SumArg& S = P.sums[si];
		std::string prod;
		for (int t = 0; t < S.nterms; t++) {
			const Term& T = S.t[t];
			snprintf(b, sizeof b, "P.sums[%d].t[%d]", si, t);
			const std::string tp = b;
			std::string term = "(" + uniform("i64", tp + ".c");
			if (T.sa)
				term += (T.sa > 0 ? " + " : " - ") + role(T.a, tp + ".a", tail_join, tail_pay);
			if (T.sb)
				term += (T.sb > 0 ? " + " : " - ") + role(T.b, tp + ".b", tail_join, tail_pay);
			term += ")";
			prod += (t ? " * (u64)" : "(u64)") + term;
		}
		return prod.empty() ? "(u64)0" : prod;
	}

	// Register accumulators of the one-group dense LDS group-by (racc)
	std::string racc_decls() const
	{
		const int ns = P.nsums > 0 ? P.nsums : 1;
		return "\tu64 racc_lo[" + std::to_string(ns) + "] = {};\n\ti64 racc_hi[" + std::to_string(ns) +
			"] = {};\n\tu32 racc_cnt = 0;\n\t(void)racc_hi;\n";
	}

	// ... added into the LDS table by every lane (before the flush's barrier)
	std::string racc_flush() const
	{
		std::string k = "\tatomicAdd(&L.cnt[0], racc_cnt);\n";
		char b[512];
		for (int si = 0; si < P.nsums; si++) {
			const SumArg& S = P.sums[si];
			if (P.wide && S.wide)
				snprintf(b, sizeof b, "\t{\n\t\tconst u64 o = atomicAdd(&L.lo[%d], racc_lo[%d]);\n"
					"\t\tconst i64 h = racc_hi[%d] + (i64)((o + racc_lo[%d]) < o);\n"
					"\t\tif (h) atomicAdd(&L.hi[%d], (int)h);\n\t}\n", si, si, si, si, si);
			else
				snprintf(b, sizeof b, "\tatomicAdd(&L.lo[%d], racc_lo[%d]);\n", si, si);
			k += b;
		}
		return k;
	}

	// Aggregate one row (u, with the last join's payload tp): group index,
	// count, sums; add_count / add_sum of vh_fused.hpp with the plan's
	// flags as literals for the dense modes.  Dense group indexes are below
	// 2^26 (the lowering's limits), so they are computed in 32 bits; the
	// last key's multiplier is 1 by construction of the mixed radix.
	std::string consume_body(int mode, int tail_join, const char* tail_pay)
	{
		std::string s;
		char b[1024];
		const std::string err = uniform("u32*", "P.error");
		const std::string dom = uniform("u64", "P.domain");
		const bool hash = mode == GB_HASH;
		s += hash ? "\t\t\tu64 g = 0;\n" : "\t\t\tu32 g = 0;\n";
		s += "\t\t\tbool ok = true;\n";
		for (int k = 0; k < P.nkeys; k++) {
			snprintf(b, sizeof b, "P.keys[%d]", k);
			const std::string kp = b;
			s += "\t\t\t{\n\t\t\t\tconst u64 o = (u64)(" + role(P.keys[k].r, kp + ".r", tail_join, tail_pay) + " - " +
				uniform("i64", kp + ".min") + ");\n";
			s += "\t\t\t\tok = ok && o <= " + uniform("u64", kp + ".range") + ";\n";
			if (hash)
				s += "\t\t\t\tg += o << " + uniform("int", kp + ".shift") + ";\n";
			else if (k == P.nkeys - 1)
				s += "\t\t\t\tg += (u32)o;\n";
			else
				s += "\t\t\t\tg += (u32)o * (u32)" + uniform("u64", kp + ".mult") + ";\n";
			s += "\t\t\t}\n";
		}
		if (hash)
			s += "\t\t\tif (!ok) { gatomic_or(" + err + ", 8u); return; }\n";
		else
			s += "\t\t\tif (!ok || (u64)g >= " + dom + ") { gatomic_or(" + err + ", 8u); return; }\n";
		if (mode == GB_PART) {
			// one tuple: a slot in the region's tuple space, then each sum's argument
			s += "\t\t\tu32 t;\n\t\t\tpart_tuple_slot(P, L, g, true, t);\n";
			for (int si = 0; si < P.nsums; si++) {
				snprintf(b, sizeof b, "\t\t\tpart_tuple_value(P, L, t, %d, g, %s);\n", si,
					sum_expr(si, tail_join, tail_pay).c_str());
				s += b;
			}
			return s;
		}
		if (mode == GB_LDS && racc) {
			// one group: the lane's register accumulators (exact: a carry
			// count beside each 64-bit sum that needs 128 bits)
			s += "\t\t\tracc_cnt++;\n";
			for (int si = 0; si < P.nsums; si++) {
				const SumArg& S = P.sums[si];
				snprintf(b, sizeof b, "\t\t\t{\n\t\t\t\tconst u64 v = %s;\n\t\t\t\tconst u64 n = racc_lo[%d] + v;\n",
					sum_expr(si, tail_join, tail_pay).c_str(), si);
				s += b;
				if (P.wide && S.wide) {
					snprintf(b, sizeof b, "\t\t\t\tracc_hi[%d] += (i64)(n < racc_lo[%d]) - (i64)(%s && (i64)v < 0);\n", si, si,
						S.arg_signed ? "true" : "false");
					s += b;
				}
				snprintf(b, sizeof b, "\t\t\t\tracc_lo[%d] = n;\n\t\t\t}\n", si);
				s += b;
			}
			return s;
		}
		const bool dense = mode == GB_REG || mode == GB_LDS || mode == GB_GLOBAL;
		if (!dense) {
			snprintf(b, sizeof b, "\t\t\tconst i64 slot = locate<%d>(P, L, g);\n", mode);
			s += b;
			if (mode == GB_HASH)
				s += "\t\t\tif (slot < 0) return;\n";
			snprintf(b, sizeof b, "\t\t\tadd_count<%d>(P, L, g, slot);\n", mode);
			s += b;
		} else if (mode == GB_REG) {
			s += "\t\t\tL.cnt[g * blockDim.x + threadIdx.x] += 1u;\n";
		} else if (mode == GB_LDS) {
			s += "\t\t\tatomicAdd(&L.cnt[g], 1u);\n";
		} else {
			s += "\t\t\tgatomic_add(&" + uniform("u64*", "P.out_cnt") + "[g], 1ull);\n";
		}
		for (int si = 0; si < P.nsums; si++) {
			const SumArg& S = P.sums[si];
			const std::string v = sum_expr(si, tail_join, tail_pay);
			const bool track = P.wide && S.wide;
			const bool exact = P.wide && S.res128;
			const char* sgn = S.arg_signed ? "true" : "false";
			if (!dense) {
				snprintf(b, sizeof b, "\t\t\tadd_sum<%d>(P, L, %d, g, slot, %s);\n", mode, si, v.c_str());
				s += b;
			} else if (mode == GB_REG) {
				snprintf(b, sizeof b,
					"\t\t\t{\n\t\t\t\tconst u64 v = %s;\n\t\t\t\tconst u32 idx = (g * %du + %du) * blockDim.x + threadIdx.x;\n"
					"\t\t\t\tconst u64 o = L.lo[idx];\n\t\t\t\tconst u64 n = o + v;\n\t\t\t\tL.lo[idx] = n;\n",
					v.c_str(), P.nsums, si);
				s += b;
				if (track)
					s += std::string("\t\t\t\tL.hi[idx] += (int)(n < o) - (int)(") + sgn + " && (i64)v < 0);\n";
				s += "\t\t\t}\n";
			} else if (mode == GB_LDS) {
				snprintf(b, sizeof b, "\t\t\t{\n\t\t\t\tconst u32 idx = %du * (u32)%s + g;\n"
					"\t\t\t\tadd_lds_exact(&L.lo[idx], &L.hi[idx], %s, %s, %s);\n\t\t\t}\n",
					si, dom.c_str(), v.c_str(), track ? "true" : "false", sgn);
				s += b;
			} else {
				snprintf(b, sizeof b, "\t\t\t{\n\t\t\t\tconst u64 idx = (u64)%d * %s + g;\n"
					"\t\t\t\tadd_global(&%s[idx], &%s[idx], %s, %s, %s);\n\t\t\t}\n",
					si, dom.c_str(), uniform("u64*", "P.out_lo").c_str(), uniform("i64*", "P.out_hi").c_str(),
					v.c_str(), exact ? "true" : "false", sgn);
				s += b;
			}
		}
		return s;
	}

	std::string col_ptr(int i)
	{
		const std::string ty = ctype_of(cols[i].type);
		return uniform(gptr(ty), "(" + gptr(ty) + ")" + cols[i].path);
	}

	// Loads of the vector path: pair p of the lane holds rows
	// chunk * 64U + 128p + 2 lane + {0, 1}; each instruction of a wave reads
	// a contiguous range (128 rows), streamed past the caches (nontemporal).
	void load_pairs(const std::string& chunk)
	{
		char b[512];
		line("\t\t{");
		line("\t\t\tconst i64 pb = (rb >> 1) + " + chunk + " * (32 * U) + lane;");
		for (size_t i = 0; i < cols.size(); i++) {
			if (cols[i].late)
				continue;
			const std::string ty = ctype_of(cols[i].type);
			line("#pragma unroll");
			snprintf(b, sizeof b, "\t\t\tfor (int p = 0; p < U / 2; p++) q%zu[p] = __builtin_nontemporal_load(&((%s)%s)[pb + p * 64]);",
				i, gptr("vh2_" + ty).c_str(), col_ptr((int)i).c_str());
			line(b);
		}
		line("\t\t}");
	}

	// Late columns of the vector loop: pair p of the lane is read only when
	// one of its two rows is live; a dead pair's buffer offset is out of
	// range, which returns zeros and reads no memory, so lines with no live
	// row are never fetched.
	void emit_late()
	{
		char b[512];
		for (size_t i = 0; i < cols.size(); i++) {
			if (!cols[i].late)
				continue;
			const int W = host_width(cols[i].type);
			const char* ty = ctype_of(cols[i].type);
			const char* fn = W == 1 ? "b16" : (W == 2 ? "b32" : (W == 4 ? "b64" : "b128"));
			line("#pragma unroll");
			line("\t\tfor (int p = 0; p < U / 2; p++) {");
			snprintf(b, sizeof b, "\t\t\tconst u32 vo = ((live >> (2 * p)) & 3u) ? (cb + (u32)p * 128u + 2u * (u32)lane) * %du "
				": 0x80000000u;", W);
			line(b);
			snprintf(b, sizeof b, "\t\t\tconst auto v = __builtin_amdgcn_raw_buffer_load_%s(rs%zu, vo, 0, 0);", fn, i);
			line(b);
			if (W == 1)
				snprintf(b, sizeof b, "\t\t\tc%zu[2 * p] = (%s)(v & 0xFFu); c%zu[2 * p + 1] = (%s)(v >> 8);", i, ty, i, ty);
			else if (W == 2)
				snprintf(b, sizeof b, "\t\t\tc%zu[2 * p] = (%s)(v & 0xFFFFu); c%zu[2 * p + 1] = (%s)(v >> 16);", i, ty, i, ty);
			else if (W == 4)
				snprintf(b, sizeof b, "\t\t\tc%zu[2 * p] = (%s)v[0]; c%zu[2 * p + 1] = (%s)v[1];", i, ty, i, ty);
			else
				snprintf(b, sizeof b, "\t\t\tc%zu[2 * p] = (i64)(((u64)v[1] << 32) | v[0]); "
					"c%zu[2 * p + 1] = (i64)(((u64)v[3] << 32) | v[2]);", i, i);
			line(b);
			line("\t\t}");
		}
	}

	// Predicates, joins and the aggregation of the rows of one chunk whose
	// columns c<i>[U] are loaded (vector path) or are loaded here (scalar
	// path: c<i>[u] = column[row[u]] for live rows, row 0 otherwise).
	// `next` requests the next chunk's columns (vector path).
	void body(int mode, bool scalar)
	{
		char b[1024];
		const int last = P.njoins - 1;
		const bool tail = P.njoins > 0 && P.joins[last].variant == VH_JOINN;
		const std::string err = uniform("u32*", "P.error");
		bool next_done = !( !scalar && pipe);
		auto emit_next = [&]() {
			if (next_done)
				return;
			next_done = true;
			// A wave's last iteration loads chunk 0 instead, which every
			// wave reads then, so those loads share lines in L2 instead of
			// re-reading each wave's own chunk from memory; no branch, which
			// would make the compiler wait for all loads where paths join.
			line("\t\t__builtin_amdgcn_sched_barrier(0);");
			load_pairs("(w + nwv < nfull ? w + nwv : 0)");
			line("\t\t__builtin_amdgcn_sched_barrier(0);");
		};
		auto load = [&](int i) {
			if (!scalar)
				return;   // loaded before the body
			const std::string ty = ctype_of(cols[i].type);
			snprintf(b, sizeof b, "\t\t%s c%d[U];", ty.c_str(), i);
			line(b);
			line("#pragma unroll");
			snprintf(b, sizeof b, "\t\tfor (int u = 0; u < U; u++) c%d[u] = %s[VH_LIVE(u) ? row[u] : 0];", i,
				col_ptr(i).c_str());
			line(b);
		};
		for (size_t i = 0; i < cols.size(); i++)
			if (cols[i].pred)
				load((int)i);
		if (eager || !scalar)
			for (size_t i = 0; i < cols.size(); i++)
				if (!cols[i].pred)
					load((int)i);
		for (int p = 0; p < P.npreds; p++) {
			snprintf(b, sizeof b, "P.preds[%d].r", p);
			const std::string v = role(P.preds[p].r, b);
			snprintf(b, sizeof b, "P.preds[%d].value", p);
			const std::string val = uniform("i64", b);
			line("#pragma unroll");
			snprintf(b, sizeof b, "\t\tfor (int u = 0; u < U; u++) VH_KILL(u, !compare(%d, %d, %s, %s));",
				P.preds[p].cmp, P.preds[p].r.type, v.c_str(), val.c_str());
			line(b);
		}
		if (!eager && scalar)
			for (size_t i = 0; i < cols.size(); i++)
				if (!cols[i].pred)
					load((int)i);
		if (!scalar && late_stage == 1 && P.njoins == 0)
			emit_late();
		if (P.njoins == 0)
			emit_next();

		// joins
		for (int j = 0; j < P.njoins; j++) {
			const JoinT& J = P.joins[j];
			snprintf(b, sizeof b, "P.joins[%d]", j);
			const std::string jp = b;
			const std::string kmin = uniform("i64", jp + ".kmin"), kmax = uniform("i64", jp + ".kmax");
			const std::string mask = uniform("u64", jp + ".mask");
			const std::string slots = uniform(gptr("u64"), "(" + gptr("u64") + ")" + jp + ".slots");
			if (!scalar && late_stage == j + 1)
				emit_late();
			snprintf(b, sizeof b, "\t\tu32 pay%d[U];", j);
			line(b);
			line("\t\t{");
			line("\t\t\tu32 off[U];");
			line("#pragma unroll");
			line("\t\t\tfor (int u = 0; u < U; u++) {");
			if (J.nkeys <= 1) {
				line("\t\t\t\tconst i64 key = " + role(J.key, jp + ".key") + ";");
				line("\t\t\t\tVH_KILL(u, key < " + kmin + " || key > " + kmax + ");");
				line("\t\t\t\toff[u] = VH_LIVE(u) ? (u32)(key - " + kmin + ") : 0u;");
			} else {
				// multi-column key: each column's offset at its bit shift
				line("\t\t\t\tu32 o = 0;");
				for (int c = 0; c < J.nkeys; c++) {
					snprintf(b, sizeof b, "%s.keyc[%d]", jp.c_str(), c);
					const std::string kp = b;
					snprintf(b, sizeof b, "%s.kminc[%d]", jp.c_str(), c);
					const std::string cmin = uniform("i64", b);
					snprintf(b, sizeof b, "%s.krangec[%d]", jp.c_str(), c);
					const std::string crange = uniform("u64", b);
					snprintf(b, sizeof b, "%s.kshiftc[%d]", jp.c_str(), c);
					const std::string csh = uniform("int", b);
					line("\t\t\t\t{");
					line("\t\t\t\t\tconst u64 d = (u64)(" + role(J.keyc[c], kp) + " - " + cmin + ");");
					line("\t\t\t\t\tVH_KILL(u, d > " + crange + ");");
					line("\t\t\t\t\to |= (u32)d << " + csh + ";");
					line("\t\t\t\t}");
				}
				line("\t\t\t\toff[u] = VH_LIVE(u) ? o : 0u;");
			}
			line("\t\t\t}");
			if (!scalar && j == last && late_stage == P.njoins + 1)
				emit_late();   // aggregation columns, before this join's lookups (and the next chunk's loads)
			if (radix && j == last) {
				// radix-partitioned probe: the row's tuple goes to the
				// workgroup's LDS staging, counted by partition; the chunk's
				// flush stores it in its partition's run, and the join kernel
				// does the lookup and the group-by
				const std::string rps = uniform("int", "P.rpshift");
				const auto vals = radix_values();
				line("#pragma unroll");
				line("\t\t\tfor (int u = 0; u < U; u++) {");
				line("\t\t\t\tif (!VH_LIVE(u)) continue;");
				line("\t\t\t\tconst u32 pos = wave_append(&rmeta[0]);");
				line("\t\t\t\tif (pos >= CRr) { gatomic_or(" + err + ", 64u); continue; }");
				line("\t\t\t\tatomicAdd(&rc[(u32)((voila_hash((u64)off[u]) & " + mask + ") >> " + rps + ")], 1u);");
				line("\t\t\t\tskey[pos] = off[u];");
				for (size_t k = 0; k < vals.size(); k++) {
					if (vals[k].first.kind == ROLE_PAY)
						snprintf(b, sizeof b, "\t\t\t\tsv%zu[pos] = pay%d[u];", k, vals[k].first.slot);
					else
						snprintf(b, sizeof b, "\t\t\t\tsv%zu[pos] = (%s)%s;", k, radix_vtype(vals[k].first).c_str(),
							role(vals[k].first, vals[k].second).c_str());
					line(b);
				}
				line("\t\t\t}");
				line("\t\t}");
				continue;
			}
			if (J.dir_w) {
				const char* et = J.dir_w == 1 ? "uint8_t" : (J.dir_w == 2 ? "uint16_t" : "uint32_t");
				const char* ev = J.dir_w == 1 ? "0xFFu" : (J.dir_w == 2 ? "0xFFFFu" : "0xFFFFFFFFu");
				const std::string dir = uniform(gptr(et), "(" + gptr(et) + ")" + jp + ".dir");
				if (J.lds == LDS_BITS) {
					// keys without a row end here, in LDS; dead rows then read entry 0
					line("#pragma unroll");
					snprintf(b, sizeof b, "\t\t\tfor (int u = 0; u < U; u++) VH_KILL(u, !((tb%d[off[u] >> 5] >> (off[u] & 31u)) & 1u));", j);
					line(b);
					line("#pragma unroll");
					line("\t\t\tfor (int u = 0; u < U; u++) off[u] = VH_LIVE(u) ? off[u] : 0u;");
				}
				line("#pragma unroll");
				if (J.lds == LDS_DIR)
					snprintf(b, sizeof b, "\t\t\tfor (int u = 0; u < U; u++) pay%d[u] = (u32)td%d[off[u]];", j, j);
				else
					snprintf(b, sizeof b, "\t\t\tfor (int u = 0; u < U; u++) pay%d[u] = (u32)%s[off[u]];", j, dir.c_str());
				line(b);
				if (j == last)
					emit_next();
				line("#pragma unroll");
				snprintf(b, sizeof b, "\t\t\tfor (int u = 0; u < U; u++) VH_KILL(u, pay%d[u] == %s);", j, ev);
				line(b);
			} else {
				// next slot: within the key's sub-table for a partitioned table
				const std::string step = J.hashed
					? "hnext(h[u], " + uniform("u64", jp + ".smask") + ")" : "(h[u] + 1) & " + mask;
				line("\t\t\tu64 h[U], s[U];");
				if (J.hashed && j == last)
					line("\t\t\tu32 ovf_first = 0;");
				line("#pragma unroll");
				line("\t\t\tfor (int u = 0; u < U; u++) h[u] = voila_hash((u64)off[u]) & " + mask + ";");
				line("#pragma unroll");
				line("\t\t\tfor (int u = 0; u < U; u++) s[u] = " + slots + "[h[u]];");
				// the next slot of each sequence is requested with the first:
				// a collision then costs no second round trip (the two slots
				// share a line but in one case of 16)
				line("\t\t\tu64 s2[U];");
				line("#pragma unroll");
				line("\t\t\tfor (int u = 0; u < U; u++) { const u64 h2 = " + std::string(J.hashed ? "hnext(h[u], " +
					uniform("u64", jp + ".smask") + ")" : "(h[u] + 1) & " + mask) + "; s2[u] = " + slots + "[h2]; }");
				line("#pragma unroll");
				line("\t\t\tfor (int u = 0; u < U; u++) {");
				snprintf(b, sizeof b, "\t\t\t\tpay%d[u] = 0;", j);
				line(b);
				line("\t\t\t\tif (!VH_LIVE(u)) continue;");
				line("\t\t\t\tu64 n = 0;");
				line("\t\t\t\tif (s[u] != kEmpty && (u32)s[u] != off[u]) { n = 1; h[u] = " + step + "; s[u] = s2[u]; }");
				// a partitioned table's sub-table may be full (its other rows
				// spilled): one full cycle of it ends the walk
				const std::string lim = J.hashed ? "(" + uniform("u64", jp + ".smask") + " + 1)" : "kMaxProbe";
				line("\t\t\t\twhile (s[u] != kEmpty && (u32)s[u] != off[u] && n < " + lim + ") {");
				line("\t\t\t\t\tn++;");
				line("\t\t\t\t\th[u] = " + step + ";");
				line("\t\t\t\t\ts[u] = " + slots + "[h[u]];");
				line("\t\t\t\t}");
				line("\t\t\t\tif (n >= kMaxProbe) gatomic_or(" + err + ", 32u);");
				if (J.hashed) {
					// not in its sub-table: a full sub-table's rows are in the
					// overflow table (whole-table wrap)
					const std::string ovf = uniform(gptr("u64"), "(" + gptr("u64") + ")" + jp + ".ovf");
					snprintf(b, sizeof b, "\t\t\t\tif ((s[u] == kEmpty || (u32)s[u] != off[u]) && ((dups_mask >> %d) & 1u)) {",
						kSpillBit + j);
					line(b);
					if (j == last)
						line("\t\t\t\t\tovf_first |= 1u << u;");
					line("\t\t\t\t\th[u] = voila_hash((u64)off[u]) & " + mask + ";");
					line("\t\t\t\t\ts[u] = " + ovf + "[h[u]];");
					line("\t\t\t\t\tfor (n = 0; s[u] != kEmpty && (u32)s[u] != off[u] && n < kMaxProbe; n++) {");
					line("\t\t\t\t\t\th[u] = (h[u] + 1) & " + mask + ";");
					line("\t\t\t\t\t\ts[u] = " + ovf + "[h[u]];");
					line("\t\t\t\t\t}");
					line("\t\t\t\t\tif (n >= kMaxProbe) gatomic_or(" + err + ", 32u);");
					line("\t\t\t\t}");
				}
				line("\t\t\t\tif (s[u] == kEmpty || (u32)s[u] != off[u]) VH_KILL(u, 1);");
				snprintf(b, sizeof b, "\t\t\t\telse pay%d[u] = (u32)(s[u] >> 32);", j);
				line(b);
				line("\t\t\t}");
				if (j == last)
					emit_next();
			}
			if (j != last && J.variant == VH_JOINN) {
				// a first-match join may not fan out: reported, as the generic kernel does
				snprintf(b, sizeof b, "\t\t\tif ((dups_mask >> %d) & 1u)", j);
				line(b);
				line("#pragma unroll");
				line("\t\t\t\tfor (int u = 0; u < U; u++) if (VH_LIVE(u)) gatomic_or(" + err + ", 16u);");
			}
			if (j == last) {
				// aggregate inside the last join's scope (its tail needs off / h)
				line("\t\t\tauto consume = [&](int u, u32 tp) __attribute__((always_inline)) {");
				out += consume_body(mode, last, "tp");
				line("\t\t\t};");
				line("#pragma unroll");
				snprintf(b, sizeof b, "\t\t\tfor (int u = 0; u < U; u++) if (VH_LIVE(u)) consume(u, pay%d[u]);", j);
				line(b);
				if (tail) {
					// further matches: after the first one in the probe sequence
					// (hash layout) or in the overflow slots from voila_hash(off)
					// on (direct layout); none without a repeated build key
					snprintf(b, sizeof b, "\t\t\tif ((dups_mask >> %d) & 1u) {", j);
					line(b);
					line("#pragma unroll");
					line("\t\t\t\tfor (int u = 0; u < U; u++) {");
					line("\t\t\t\t\tif (!VH_LIVE(u)) continue;");
					if (J.hashed) {
						// the rest of the match's sub-table sequence, then
						// the overflow table if the build spilled
						const std::string sm = uniform("u64", jp + ".smask");
						const std::string ovf = uniform(gptr("u64"), "(" + gptr("u64") + ")" + jp + ".ovf");
						line("\t\t\t\t\tbool ov = (ovf_first >> u) & 1u;");
						line("\t\t\t\t\tconst u64 h0 = voila_hash((u64)off[u]) & " + mask + ";");
						line("\t\t\t\t\tu64 hh = ov ? (h[u] + 1) & " + mask + " : hnext(h[u], " + sm + ");");
						line("\t\t\t\t\tfor (u64 n = 0;; n++) {");
						line("\t\t\t\t\t\tif (n >= kMaxProbe) { gatomic_or(" + err + ", 32u); break; }");
						// a full sub-table has no empty slot: its sequence ends where it began
						line("\t\t\t\t\t\tconst u64 sv = ov ? " + ovf + "[hh] : (hh == h0 ? kEmpty : " + slots + "[hh]);");
						line("\t\t\t\t\t\tif (sv == kEmpty) {");
						snprintf(b, sizeof b, "\t\t\t\t\t\t\tif (!ov && ((dups_mask >> %d) & 1u)) {", kSpillBit + j);
						line(b);
						line("\t\t\t\t\t\t\t\tov = true;");
						line("\t\t\t\t\t\t\t\thh = voila_hash((u64)off[u]) & " + mask + ";");
						line("\t\t\t\t\t\t\t\tcontinue;");
						line("\t\t\t\t\t\t\t}");
						line("\t\t\t\t\t\t\tbreak;");
						line("\t\t\t\t\t\t}");
						line("\t\t\t\t\t\tif ((u32)sv == off[u]) consume(u, (u32)(sv >> 32));");
						line("\t\t\t\t\t\thh = ov ? (hh + 1) & " + mask + " : hnext(hh, " + sm + ");");
						line("\t\t\t\t\t}");
						line("\t\t\t\t}");
					} else {
					if (J.dir_w)
						line("\t\t\t\t\tu64 hh = voila_hash((u64)off[u]) & " + mask + ";");
					else
						line("\t\t\t\t\tu64 hh = (h[u] + 1) & " + mask + ";");
					line("\t\t\t\t\tfor (u64 n = 0;; n++) {");
					line("\t\t\t\t\t\tif (n >= kMaxProbe) { gatomic_or(" + err + ", 32u); break; }");
					line("\t\t\t\t\t\tconst u64 sv = " + slots + "[hh];");
					line("\t\t\t\t\t\tif (sv == kEmpty) break;");
					line("\t\t\t\t\t\tif ((u32)sv == off[u]) consume(u, (u32)(sv >> 32));");
					line("\t\t\t\t\t\thh = (hh + 1) & " + mask + ";");
					line("\t\t\t\t\t}");
					line("\t\t\t\t}");
					}
					line("\t\t\t}");
				}
			}
			line("\t\t}");
		}
		if (P.njoins == 0) {
			line("\t\t{");
			line("\t\t\tauto consume = [&](int u) __attribute__((always_inline)) {");
			out += consume_body(mode, -1, nullptr);
			line("\t\t\t};");
			line("#pragma unroll");
			line("\t\t\tfor (int u = 0; u < U; u++) if (VH_LIVE(u)) consume(u);");
			line("\t\t}");
		}
	}

	// Rows [first, end) of the scalar path: chunks of U * blockDim rows, row
	// base + u * blockDim + thread, every `stride` rows.
	void scalar_loop(int mode, const std::string& first, const std::string& end, const std::string& stride)
	{
		line("\tfor (i64 base = " + first + "; base < " + end + "; base += " + stride + ") {");
		line("\t\ti64 row[U];");
		line("\t\tu32 live = 0;");
		line("#pragma unroll");
		line("\t\tfor (int u = 0; u < U; u++) {");
		line("\t\t\trow[u] = base + (i64)u * blockDim.x + threadIdx.x;");
		line("\t\t\tlive |= (u32)(row[u] < " + end + ") << u;");
		line("\t\t}");
		body(mode, true);
		line("\t}");
	}

	// One wave chunk of the vector loop (64U rows from wave chunk `w`):
	// its column pairs, the chunk's first row, the rows' values, the body.
	void vec_chunk(int mode, bool declare_pairs)
	{
		char b[512];
		if (declare_pairs)
			for (size_t i = 0; i < cols.size(); i++) {
				if (cols[i].late)
					continue;
				snprintf(b, sizeof b, "\t\tvh2_%s q%zu[U / 2];", ctype_of(cols[i].type), i);
				line(b);
			}
		if (!pipe)
			load_pairs("w");
		line("\t\tconst u32 cb = (u32)(rb + w * (64 * U));   // first row of the chunk");
		line("\t\t(void)cb;");
		for (size_t i = 0; i < cols.size(); i++) {
			snprintf(b, sizeof b, "\t\t%s c%zu[U];", ctype_of(cols[i].type), i);
			line(b);
			if (cols[i].late)
				continue;
			line("#pragma unroll");
			snprintf(b, sizeof b, "\t\tfor (int p = 0; p < U / 2; p++) { c%zu[2 * p] = q%zu[p].x; c%zu[2 * p + 1] = "
				"q%zu[p].y; }", i, i, i, i);
			line(b);
		}
		snprintf(b, sizeof b, "\t\tu32 live = 0x%xu;", (unsigned)((1ull << U) - 1));
		line(b);
		body(mode, false);
	}

	// Row loops of the radix probe kernel.  Every iteration of a workgroup
	// takes one chunk of U * blockDim rows and ends with the chunk's flush,
	// so every wave of the workgroup runs the same iterations (the flush
	// has barriers).  Vector path: chunk c is wave chunks [c * waves,
	// (c + 1) * waves); the rows after the last whole wave chunk go through
	// the scalar loop of workgroup 0 as chunk nmain.  Scalar path: chunk c is
	// rows [c * CR, (c + 1) * CR).  Workgroup 0 then writes empty runs for
	// the chunk ids after the last one used.
	void radix_loops(int mode)
	{
		line("\tconst i64 CR = (i64)CRr;   // U * blockDim.x rows per chunk");
		if (vec) {
			line("\tconst int lane = threadIdx.x & 63;");
			line("\tconst i64 rb = 0;");
			line("\tconst i64 nfull = rows / (64 * U);");
			line("\tconst i64 wpb = (i64)(blockDim.x >> 6);");
			line("\tconst i64 wsub = (i64)__builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));");
			line("\tconst i64 nmain = (nfull + wpb - 1) / wpb;");
			line("\tfor (i64 ci = blockIdx.x; ci < nmain; ci += gridDim.x) {");
			line("\t\tconst i64 w = ci * wpb + wsub;");
			line("\t\tif (w < nfull) {");
			vec_chunk(mode, true);
			line("\t\t}");
			radix_flush("ci");
			line("\t}");
			line("\tif (blockIdx.x == 0) {");
			scalar_loop(mode, "nfull * (64 * U)", "rows", "tile");
			radix_flush("nmain");
			line("\t\tradix_pad(P, (u32)nmain + 1u);");
			line("\t}");
		} else {
			line("\tconst i64 nch = (rows + CR - 1) / CR;");
			line("\tfor (i64 ci = blockIdx.x; ci < nch; ci += gridDim.x) {");
			line("\t\tconst i64 ce = ci * CR + CR < rows ? ci * CR + CR : rows;");
			scalar_loop(mode, "ci * CR", "ce", "tile");
			radix_flush("ci");
			line("\t}");
			line("\tif (blockIdx.x == 0)");
			line("\t\tradix_pad(P, (u32)nch);");
		}
	}

	std::string source(int mode)
	{
		char b[1024];
		racc = mode == GB_LDS && P.domain == 1 && !radix;
		// register every probe column: predicates first, then join keys,
		// group keys and sum arguments
		for (int p = 0; p < P.npreds; p++) {
			snprintf(b, sizeof b, "P.preds[%d].r", p);
			note_role(P.preds[p].r, b, true, 0);
		}
		for (int j = 0; j < P.njoins; j++) {
			if (P.joins[j].nkeys <= 1) {
				snprintf(b, sizeof b, "P.joins[%d].key", j);
				note_role(P.joins[j].key, b, false, j + 1);
			} else {
				for (int c = 0; c < P.joins[j].nkeys; c++) {
					snprintf(b, sizeof b, "P.joins[%d].keyc[%d]", j, c);
					note_role(P.joins[j].keyc[c], b, false, j + 1);
				}
			}
		}
		for (int k = 0; k < P.nkeys; k++) {
			snprintf(b, sizeof b, "P.keys[%d].r", k);
			note_role(P.keys[k].r, b, false, P.njoins + 1);
		}
		for (int si = 0; si < P.nsums; si++)
			for (int t = 0; t < P.sums[si].nterms; t++) {
				const Term& T = P.sums[si].t[t];
				snprintf(b, sizeof b, "P.sums[%d].t[%d]", si, t);
				if (T.sa)
					note_role(T.a, std::string(b) + ".a", false, P.njoins + 1);
				if (T.sb)
					note_role(T.b, std::string(b) + ".b", false, P.njoins + 1);
			}
		// The vector main loop needs every probe column aligned to a pair of
		// its values and an even U; otherwise every row takes the scalar path.
		vec = U % 2 == 0;
		if (radix)
			pipe = false;   // the last join's lookups are in the join kernel
		int row_bytes = 0;
		for (const Col& c : cols) {
			vec = vec && ((uintptr_t)c.ptr % (uintptr_t)(2 * host_width(c.type))) == 0;
			row_bytes += host_width(c.type);
		}
		bool hash_join = false;
		for (int j = 0; j < P.njoins; j++)
			hash_join = hash_join || !P.joins[j].dir_w;
		if (U == 0) {
			// measured on MI355X: 8 rows per lane for 16 (c1 probe) and 32
			// bytes (star fact) of columns per row, 2 for 36 (filter); 4 in
			// 1024-thread workgroups (star: 344 us against 351 with 8) and
			// with hash-table lookups, which keep more registers per row
			// (sparse-key join: 86 us at 4 against 98 at 8)
			U = vec ? (block > 256 || hash_join ? 4 : (row_bytes <= 32 ? 8 : 2)) : 4;
		}
		// Late materialization in the vector loop: columns first used at or
		// after late_stage are loaded there, for live rows only.  Asked for
		// (not eager): after the last join whose build is filtered, else
		// after the predicates.  By default, when two joins have filtered
		// builds (a dimension filter keeping 1 in 5 twice leaves 4% of the
		// rows, and about half of the 8-byte lines hold no live row): the
		// aggregation's columns only, requested before the last join's
		// lookups, since a late join key costs an exposed memory round trip
		// (measured: star probe 465 us with late date/part keys, 458 us
		// without late columns).
		if (vec) {
			int last_filtered = -1, nfiltered = 0;
			for (int j = 0; j < P.njoins; j++)
				if (P.joins[j].filtered) {
					last_filtered = j;
					nfiltered++;
				}
			if (!eager)
				late_stage = last_filtered >= 0 ? last_filtered + 2 : (P.npreds > 0 ? 1 : 0);
			else if (nfiltered >= 2)
				late_stage = P.njoins + 1;   // aggregation columns only: their loads
				                             // overlap the last join's lookups
			for (Col& c : cols)
				c.late = late_stage > 0 && c.stage >= late_stage &&
					(double)P.rows * host_width(c.type) < 2147483000.0;   // buffer offsets are 32-bit
		}

		// rows of this workgroup only (local): a batch item, or GB_PART's
		// slab, the region its tuples are kept for; else grid-strided
		bool local = batch;
		std::string lrb = "I.r0", lre = "I.r1";
		if (!radix && mode == GB_PART) {
			if (batch) {
				line("\tconst i64 first = I.r0;");
				line("\tconst i64 end = I.r1;");
				line("\tL.region = (u32)(first / P.gregion_rows);");
			} else {
				line("\tconst i64 first = (i64)blockIdx.x * P.gregion_rows;");
				line("\tconst i64 end = first + P.gregion_rows < rows ? first + P.gregion_rows : rows;");
				line("\tL.region = blockIdx.x;");
			}
			line("\tL.rbase = first;");
			line("\tL.rcap = end > first ? (u32)(end - first) : 0u;");
			local = true;
			lrb = "first";
			lre = "end";
		}
		if (radix) {
			radix_loops(mode);
		} else if (!vec) {
			if (local)
				scalar_loop(mode, lrb, lre, "tile");
			else
				scalar_loop(mode, "(i64)blockIdx.x * tile", "rows", "(i64)gridDim.x * tile");
		} else {
			// full chunks of 64U rows of [rb, re), one per wave at a time, then
			// the remaining rows (fewer than 64U) on the scalar path.  A
			// workgroup's own rows go to its waves (rb is even: items and
			// GB_PART slabs start at multiples of 1024 rows)
			line("\tconst int lane = threadIdx.x & 63;");
			if (local) {
				line("\tconst i64 rb = " + lrb + ", re = " + lre + ";");
				line("\tconst i64 wv = (i64)__builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));");
				line("\tconst i64 nwv = (i64)(blockDim.x >> 6);");
			} else {
				line("\tconst i64 rb = 0, re = rows;");
				line("\tconst i64 wv = (i64)__builtin_amdgcn_readfirstlane((int)(blockIdx.x * (blockDim.x >> 6) + "
					"(threadIdx.x >> 6)));");
				line("\tconst i64 nwv = (i64)gridDim.x * (blockDim.x >> 6);");
			}
			line("\tconst i64 nfull = (re - rb) / (64 * U);");
			line("\tif (wv < nfull) {");
			for (size_t i = 0; i < cols.size(); i++) {
				if (cols[i].late)
					continue;
				snprintf(b, sizeof b, "\t\tvh2_%s q%zu[U / 2];", ctype_of(cols[i].type), i);
				line(b);
			}
			// pipelined: the first chunk here, each next one behind the last
			// join's lookups (body); otherwise every chunk at the loop top
			if (pipe)
				load_pairs("wv");
			line("\tfor (i64 w = wv; w < nfull; w += nwv) {");
			vec_chunk(mode, false);
			line("\t}");
			line("\t}");
			if (local)
				scalar_loop(mode, "rb + nfull * (64 * U)", "re", "tile");
			else
				scalar_loop(mode, "rb + nfull * (64 * U) + (i64)blockIdx.x * tile", "re", "(i64)gridDim.x * tile");
		}
		const std::string loop = out;

		// ---- assemble ----
		std::string k;
		k += "// generated probe kernel\n";
		for (const char* t : {"int8_t", "uint8_t", "int16_t", "uint16_t", "int32_t", "uint32_t", "int64_t", "uint64_t"})
			k += std::string("typedef __hip_internal::") + t + " " + t + ";\n";
		k += "#include \"vh_fused.hpp\"\nusing namespace vh;\n";
		for (const char* t : {"int8_t", "uint8_t", "int16_t", "uint16_t", "int32_t", "uint32_t", "i64"})
			k += std::string("typedef ") + t + " vh2_" + t + " __attribute__((ext_vector_type(2)));\n";
		k += "#define VH_LIVE(u) ((live >> (u)) & 1u)\n";
		k += "#define VH_KILL(u, c) (live &= ~((u32)(c) << (u)))\n";
		k += "static constexpr int U = " + std::to_string(U) + ";\n";
		if (batch) {
			// persistent: work items from the launch's queue until it is empty
			k += "extern \"C\" __global__ __launch_bounds__(" + std::to_string(block) +
				") void vh_generated_batch(const Plan* __restrict__ plans, const BatchItem* __restrict__ items, "
				"u32 first_item, u32 count, u32* __restrict__ counter)\n{\n";
			k += "\textern __shared__ u64 lds[];\n\t__shared__ u32 slot;\n";
			k += "\tfor (;;) {\n";
			k += "\tif (threadIdx.x == 0)\n\t\tslot = gatomic_add(counter, 1u);\n\t__syncthreads();\n";
			k += "\tconst u32 item = slot;\n\t__syncthreads();\n\tif (item >= count)\n\t\tbreak;\n";
			k += "\tconst BatchItem I = items[first_item + item];\n";
			k += "\tconst Plan& P = plans[I.plan];\n";
		} else {
			k += "extern \"C\" __global__ __launch_bounds__(" + std::to_string(block) +
				") void vh_generated(const Plan* __restrict__ Pp)\n{\n";
			k += "\tconst Plan& P = *Pp;\n\textern __shared__ u64 lds[];\n";
		}
		if (radix) {
			radix_stage_decls(k);
		} else {
			k += "\tLds L = lds_layout<" + std::to_string(mode) + ">(P, lds);\n";
			k += "\tlds_init<" + std::to_string(mode) + ">(P, L);\n";
		}
		// join tables kept in LDS: copied by every workgroup (16-byte units;
		// dir[] and the bitmap are allocated in whole units)
		for (int j = 0; j < P.njoins; j++) {
			const JoinT& J = P.joins[j];
			if (!J.lds)
				continue;
			const std::string jp = "P.joins[" + std::to_string(j) + "]";
			const char* et = J.dir_w == 1 ? "uint8_t" : (J.dir_w == 2 ? "uint16_t" : "uint32_t");
			const std::string keys = "((u64)(" + jp + ".kmax - " + jp + ".kmin) + 1)";
			const std::string bytes = J.lds == LDS_DIR ? keys + " * " + std::to_string(J.dir_w) + "u"
				: "(" + keys + " + 31) / 32 * 4";
			const std::string src = J.lds == LDS_DIR ? jp + ".dir" : jp + ".bits";
			k += "\t{\n\t\tconst u64 n = (" + bytes + " + 15) / 16;\n";
			k += "\t\tu64* dst = (u64*)((char*)lds + " + jp + ".lds_off);\n";
			k += "\t\tconst u64* src = (const u64*)" + src + ";\n";
			k += "\t\tfor (u64 i = threadIdx.x; i < n; i += blockDim.x) { dst[2 * i] = gload(src, (i64)(2 * i)); "
				"dst[2 * i + 1] = gload(src, (i64)(2 * i + 1)); }\n\t}\n";
			if (J.lds == LDS_DIR)
				k += "\tconst " + std::string(et) + "* const td" + std::to_string(j) + " = (const " + et +
					"*)((const char*)lds + " + jp + ".lds_off);\n";
			else
				k += "\tconst u32* const tb" + std::to_string(j) + " = (const u32*)((const char*)lds + " + jp +
					".lds_off);\n";
		}
		k += "\tu32 dups_mask = 0;\n";
		for (int j = 0; j < P.njoins; j++) {
			if (P.joins[j].variant == VH_JOINN)
				k += "\tif (gload(P.joins[" + std::to_string(j) + "].dups, 0)) dups_mask |= " + std::to_string(1u << j) + "u;\n";
			if (P.joins[j].hashed)
				k += "\tif (gload(P.joins[" + std::to_string(j) + "].spill, 0)) dups_mask |= " +
					std::to_string(1u << (kSpillBit + j)) + "u;\n";
		}
		k += "\t(void)dups_mask;\n";
		k += decls;
		for (size_t i = 0; i < cols.size(); i++)
			if (cols[i].late)
				k += "\tconst __amdgpu_buffer_rsrc_t rs" + std::to_string(i) +
					" = __builtin_amdgcn_make_buffer_rsrc((void*)" + cols[i].path + ", (short)0, (int)(P.rows * " +
					std::to_string(host_width(cols[i].type)) + "), 0x00020000);\n";
		if (racc)
			k += racc_decls();
		k += "\t__syncthreads();\n\tconst i64 rows = P.rows;\n\tconst i64 tile = (i64)U * blockDim.x;\n";
		k += loop;
		if (racc)
			k += racc_flush();
		if (batch) {
			// the item's LDS group-by state goes to the pipeline's global
			// arrays before the next item reuses the LDS
			k += "\t__syncthreads();\n";
			if (mode == GB_LDS)
				k += "\tlds_dense_flush(P, L);\n";
			else if (mode == GB_LDSHASH)
				k += "\tldshash_flush(P, L);\n";
			else if (mode == GB_PART)
				k += "\tpart_finish(P, L);\n";
			k += "\t__syncthreads();\n\t}\n}\n";
		} else if (radix) {
			k += "}\n";
		} else {
			k += "\tblock_epilogue<" + std::to_string(mode) + ">(P, L);\n}\n";
		}
		return k;
	}

	// The join kernel of a radix-partitioned probe: workgroup p copies
	// partition p's 2^rpshift slots of the last join's table (whole
	// sub-tables of its partitioned build) into LDS, then takes the
	// partition's tuples (U per lane, from every chunk's run), looks each key
	// up in LDS (rows of a full sub-table: in the overflow table) and
	// aggregates it as the one-pass kernel would; at its end the LDS
	// group-by state is added to the global arrays.
	std::string join_source(int mode)
	{
		char b[1024];
		racc = mode == GB_LDS && P.domain == 1;
		const int last = P.njoins - 1;
		const JoinT& J = P.joins[last];
		snprintf(b, sizeof b, "P.joins[%d]", last);
		const std::string jp = b;
		const std::string err = uniform("u32*", "P.error");
		const std::string mask = uniform("u64", jp + ".mask");
		const std::string sm = uniform("u64", jp + ".smask");
		const std::string pm = uniform("u64", "((1ull << P.rpshift) - 1)");
		const std::string ovf = uniform(gptr("u64"), "(" + gptr("u64") + ")" + jp + ".ovf");
		const auto vals = radix_values();
		if (U <= 0)
			U = 4;
		std::vector<int> ci(vals.size(), -1);
		for (size_t k = 0; k < vals.size(); k++)
			if (vals[k].first.kind == ROLE_COL)
				ci[k] = col_index(vals[k].first, vals[k].second, false);
		line("\tfor (u32 i0 = 0; i0 < rtotal; i0 += U * blockDim.x) {");
		line("\t\tu32 live = 0;");
		line("\t\tu32 off[U];");
		for (size_t k = 0; k < vals.size(); k++) {
			if (vals[k].first.kind == ROLE_PAY)
				snprintf(b, sizeof b, "\t\tu32 pay%d[U];", vals[k].first.slot);
			else
				snprintf(b, sizeof b, "\t\t%s c%d[U];", ctype_of(vals[k].first.type), ci[k]);
			line(b);
		}
		line("#pragma unroll");
		line("\t\tfor (int u = 0; u < U; u++) {");
		line("\t\t\tconst u32 i = i0 + u * blockDim.x + threadIdx.x;");
		line("\t\t\tlive |= (u32)(i < rtotal) << u;");
		line("\t\t\tconst u64 x = i < rtotal ? radix_tuple_index(rpre, rbase, R, i) : 0ull;");
		line("\t\t\toff[u] = rk[x];");
		for (size_t k = 0; k < vals.size(); k++) {
			if (vals[k].first.kind == ROLE_PAY)
				snprintf(b, sizeof b, "\t\t\tpay%d[u] = rv%zu[x];", vals[k].first.slot, k);
			else
				snprintf(b, sizeof b, "\t\t\tc%d[u] = rv%zu[x];", ci[k], k);
			line(b);
		}
		line("\t\t}");
		// lookups in the LDS copy (local slot = global slot & pm; the
		// sequence wraps within the build's sub-table of smask + 1 slots)
		line("\t\tu32 payl[U];");
		line("\t\tu64 h[U], s[U];");
		line("\t\tu32 ovf_first = 0;");
		line("#pragma unroll");
		line("\t\tfor (int u = 0; u < U; u++) h[u] = voila_hash((u64)off[u]) & " + pm + ";");
		line("#pragma unroll");
		line("\t\tfor (int u = 0; u < U; u++) s[u] = tab[h[u]];");
		line("#pragma unroll");
		line("\t\tfor (int u = 0; u < U; u++) {");
		line("\t\t\tpayl[u] = 0;");
		line("\t\t\tif (!VH_LIVE(u)) continue;");
		line("\t\t\tu64 n = 0;");
		line("\t\t\twhile (s[u] != kEmpty && (u32)s[u] != off[u] && n <= " + sm + ") {");
		line("\t\t\t\tn++;");
		line("\t\t\t\th[u] = hnext(h[u], " + sm + ");");
		line("\t\t\t\ts[u] = tab[h[u]];");
		line("\t\t\t}");
		snprintf(b, sizeof b, "\t\t\tif ((s[u] == kEmpty || (u32)s[u] != off[u]) && ((dups_mask >> %d) & 1u)) {", kSpillBit + last);
		line(b);
		line("\t\t\t\tovf_first |= 1u << u;");
		line("\t\t\t\th[u] = voila_hash((u64)off[u]) & " + mask + ";");
		line("\t\t\t\ts[u] = " + ovf + "[h[u]];");
		line("\t\t\t\tfor (n = 0; s[u] != kEmpty && (u32)s[u] != off[u] && n < kMaxProbe; n++) {");
		line("\t\t\t\t\th[u] = (h[u] + 1) & " + mask + ";");
		line("\t\t\t\t\ts[u] = " + ovf + "[h[u]];");
		line("\t\t\t\t}");
		line("\t\t\t\tif (n >= kMaxProbe) gatomic_or(" + err + ", 32u);");
		line("\t\t\t}");
		line("\t\t\tif (s[u] == kEmpty || (u32)s[u] != off[u]) VH_KILL(u, 1);");
		line("\t\t\telse payl[u] = (u32)(s[u] >> 32);");
		line("\t\t}");
		line("\t\tauto consume = [&](int u, u32 tp) __attribute__((always_inline)) {");
		out += consume_body(mode, last, "tp");
		line("\t\t};");
		line("#pragma unroll");
		line("\t\tfor (int u = 0; u < U; u++) if (VH_LIVE(u)) consume(u, payl[u]);");
		if (J.variant == VH_JOINN) {
			// further matches: the rest of the sub-table's sequence (it ends
			// where it began when the sub-table is full), then the overflow
			// table after a spill
			snprintf(b, sizeof b, "\t\tif ((dups_mask >> %d) & 1u) {", last);
			line(b);
			line("#pragma unroll");
			line("\t\t\tfor (int u = 0; u < U; u++) {");
			line("\t\t\t\tif (!VH_LIVE(u)) continue;");
			line("\t\t\t\tbool ov = (ovf_first >> u) & 1u;");
			line("\t\t\t\tconst u64 h0 = voila_hash((u64)off[u]) & " + pm + ";");
			line("\t\t\t\tu64 hh = ov ? (h[u] + 1) & " + mask + " : hnext(h[u], " + sm + ");");
			line("\t\t\t\tfor (u64 n = 0;; n++) {");
			line("\t\t\t\t\tif (n >= kMaxProbe) { gatomic_or(" + err + ", 32u); break; }");
			line("\t\t\t\t\tconst u64 sv = ov ? " + ovf + "[hh] : (hh == h0 ? kEmpty : tab[hh]);");
			line("\t\t\t\t\tif (sv == kEmpty) {");
			snprintf(b, sizeof b, "\t\t\t\t\t\tif (!ov && ((dups_mask >> %d) & 1u)) {", kSpillBit + last);
			line(b);
			line("\t\t\t\t\t\t\tov = true;");
			line("\t\t\t\t\t\t\thh = voila_hash((u64)off[u]) & " + mask + ";");
			line("\t\t\t\t\t\t\tcontinue;");
			line("\t\t\t\t\t\t}");
			line("\t\t\t\t\t\tbreak;");
			line("\t\t\t\t\t}");
			line("\t\t\t\t\tif ((u32)sv == off[u]) consume(u, (u32)(sv >> 32));");
			line("\t\t\t\t\thh = ov ? (hh + 1) & " + mask + " : hnext(hh, " + sm + ");");
			line("\t\t\t\t}");
			line("\t\t\t}");
			line("\t\t}");
		}
		line("\t}");
		const std::string loop = out;

		std::string k;
		k += "// generated join kernel of a radix-partitioned probe\n";
		for (const char* t : {"int8_t", "uint8_t", "int16_t", "uint16_t", "int32_t", "uint32_t", "int64_t", "uint64_t"})
			k += std::string("typedef __hip_internal::") + t + " " + t + ";\n";
		k += "#include \"vh_fused.hpp\"\nusing namespace vh;\n";
		k += "typedef u32 vh_u32x4 __attribute__((ext_vector_type(4)));\n";
		k += "#define VH_LIVE(u) ((live >> (u)) & 1u)\n";
		k += "#define VH_KILL(u, c) (live &= ~((u32)(c) << (u)))\n";
		k += "static constexpr int U = " + std::to_string(U) + ";\n";
		k += "extern \"C\" __global__ __launch_bounds__(" + std::to_string(block) +
			") void vh_generated_join(const Plan* __restrict__ Pp)\n{\n";
		k += "\tconst Plan& P = *Pp;\n\textern __shared__ u64 lds[];\n";
		k += "\tLds L = lds_layout<" + std::to_string(mode) + ">(P, lds);\n";
		k += "\tlds_init<" + std::to_string(mode) + ">(P, L);\n";
		k += "\tconst u32 S = 1u << P.rpshift;\n";
		k += "\tu64* const tab = (u64*)((char*)lds + P.rlds_b);\n";
		k += "\tu32* const rpre = (u32*)(tab + S);\n\tu32* const rbase = rpre + P.rchunks;\n";
		k += "\tu32* const rtot = rbase + P.rchunks;\n";
		k += "\tconst int R = P.rchunks;\n";
		// the partition's slots, eight 16-byte loads in flight per thread
		k += "\t{\n\t\t" + gptr("vh_u32x4") + " src = (" + gptr("vh_u32x4") + ")(" + jp + ".slots + (u64)blockIdx.x * S);\n";
		k += "\t\tvh_u32x4* dst = (vh_u32x4*)tab;\n";
		k += "\t\tfor (u32 i0 = 0; i0 < S / 2; i0 += 8 * blockDim.x) {\n";
		k += "\t\t\tvh_u32x4 v_[8];\n";
		k += "#pragma unroll\n\t\t\tfor (int j = 0; j < 8; j++) { const u32 i = i0 + j * blockDim.x + threadIdx.x; "
			"if (i < S / 2) v_[j] = src[i]; }\n";
		k += "#pragma unroll\n\t\t\tfor (int j = 0; j < 8; j++) { const u32 i = i0 + j * blockDim.x + threadIdx.x; "
			"if (i < S / 2) dst[i] = v_[j]; }\n";
		k += "\t\t}\n\t}\n";
		k += "\tconst u32 rtotal = radix_runs(P, blockIdx.x, rpre, rbase, rtot);\n";
		k += "\t" + gptr("u32") + " const rk = (" + gptr("u32") + ")P.rkey;\n";
		for (size_t i = 0; i < vals.size(); i++) {
			const std::string t = radix_vtype(vals[i].first);
			k += "\t" + gptr(t) + " const rv" + std::to_string(i) + " = (" + gptr(t) + ")P.rval[" +
				std::to_string(i) + "];\n";
		}
		k += "\tu32 dups_mask = 0;\n";
		if (J.variant == VH_JOINN)
			k += "\tif (gload(" + jp + ".dups, 0)) dups_mask |= " + std::to_string(1u << last) + "u;\n";
		k += "\tif (gload(" + jp + ".spill, 0)) dups_mask |= " + std::to_string(1u << (kSpillBit + last)) + "u;\n";
		k += decls;
		if (racc)
			k += racc_decls();
		k += "\t__syncthreads();\n";
		k += loop;
		if (racc)
			k += racc_flush();
		k += "\t__syncthreads();\n";
		if (mode == GB_LDS)
			k += "\tlds_dense_flush(P, L);\n";
		else if (mode == GB_LDSHASH)
			k += "\tldshash_flush(P, L);\n";
		k += "}\n";
		return k;
	}
};

struct Compiled {
	hipModule_t mod = nullptr;
	hipFunction_t fn = nullptr;
};

std::mutex g_mu;
std::map<std::string, Compiled> g_cache;

std::string device_arch()
{
	int dev = 0;
	hipDeviceProp_t prop;
	if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess)
		return "";
	std::string a = prop.gcnArchName;
	const size_t c = a.find(':');
	return c == std::string::npos ? a : a.substr(0, c);
}

} // namespace

std::string jit_source(const Plan& P, bool eager, int unroll, bool pipeline, int block, bool batch)
{
	Gen g(P, eager, pipeline, unroll);
	g.block = block;
	g.batch = batch;
	g.radix = P.radix && !batch;
	return g.source(P.mode);
}

std::string jit_join_source(const Plan& P, int unroll, int block)
{
	Gen g(P, true, false, unroll);
	g.block = block;
	return g.join_source(P.mode);
}

int jit_radix_values(const Plan& P, std::vector<int>& widths)
{
	Gen g(P, true, false, 0);
	widths.clear();
	int bytes = 4;
	for (const auto& v : g.radix_values()) {
		widths.push_back(Gen::radix_vwidth(v.first));
		bytes += widths.back();
	}
	return bytes;
}

size_t jit_radix_stage_bytes(const Plan& P, u64 chunk)
{
	Gen g(P, true, false, 0);
	return g.radix_stage_bytes(chunk);
}

int jit_compile(const std::string& src, hipFunction_t* fn, std::string& log, const char* name)
{
	const std::string arch = device_arch();
	if (arch.empty())
		return fail(VH_ERR_NODEVICE, "no device to generate a kernel for");
	const std::string key = arch + "\n" + src;
	std::lock_guard<std::mutex> lock(g_mu);
	auto it = g_cache.find(key);
	if (it != g_cache.end()) {
		*fn = it->second.fn;
		return VH_OK;
	}
	hiprtcProgram prog;
	if (hiprtcCreateProgram(&prog, src.c_str(), "vh_generated.hip", kJitHeaders, kJitHeaderText, kJitHeaderNames) !=
		HIPRTC_SUCCESS)
		return fail(VH_ERR_HIP, "hiprtcCreateProgram failed");
	const std::string archopt = "--offload-arch=" + arch;
	const char* opts[] = {archopt.c_str(), "-O3", "-std=c++17", "-Xclang", "-target-feature", "-Xclang",
		"-packed-fp32-ops"};
	const hiprtcResult rc = hiprtcCompileProgram(prog, (int)(sizeof(opts) / sizeof(opts[0])), opts);
	size_t n = 0;
	hiprtcGetProgramLogSize(prog, &n);
	log.assign(n, '\0');
	if (n)
		hiprtcGetProgramLog(prog, &log[0]);
	if (rc != HIPRTC_SUCCESS) {
		hiprtcDestroyProgram(&prog);
		return fail(VH_ERR_UNSUPPORTED, "hipRTC: %s: %.400s", hiprtcGetErrorString(rc), log.c_str());
	}
	size_t sz = 0;
	hiprtcGetCodeSize(prog, &sz);
	std::vector<char> code(sz);
	hiprtcGetCode(prog, code.data());
	hiprtcDestroyProgram(&prog);
	Compiled c;
	hipError_t e = hipModuleLoadData(&c.mod, code.data());
	if (e == hipSuccess)
		e = hipModuleGetFunction(&c.fn, c.mod, name);
	if (e != hipSuccess)
		return fail(VH_ERR_HIP, "loading the generated kernel: %s", hipGetErrorString(e));
	g_cache[key] = c;
	*fn = c.fn;
	return VH_OK;
}

hipError_t launch_generated(hipFunction_t fn, const Plan& P, const Plan* dP, int grid, int block, hipStream_t s,
	hipEvent_t start, hipEvent_t stop, long lds_bytes)
{
	const Plan* arg = dP;
	void* args[] = {(void*)&arg};
	const unsigned lds = lds_bytes >= 0 ? (unsigned)lds_bytes : (unsigned)pipeline_lds_bytes(P);
	if (start || stop)
		return hipExtModuleLaunchKernel(fn, (uint32_t)grid * (uint32_t)block, 1, 1, (uint32_t)block, 1, 1, lds, s,
			args, nullptr, start, stop, 0);
	return hipModuleLaunchKernel(fn, (unsigned)grid, 1, 1, (unsigned)block, 1, 1, lds, s, args, nullptr);
}

} // namespace vh
