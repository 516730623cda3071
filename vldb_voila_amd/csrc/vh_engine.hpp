// Host-side engine state shared by the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/voila_hip.h"
#include "vh_plan.hpp"

namespace vh {

struct TypeProps {     // ScalarTypeProps (typing_pass.cpp) + exact value interval
	int type = VH_INVALID;
	double dmin = 0, dmax = 0;   // range the typing rules derive
	double emin = 0, emax = 0;   // exact interval of the value
};

struct Typing {
	std::vector<TypeProps> vcols;
	std::vector<TypeProps> nodes;
	std::vector<TypeProps> aggrs;        // first-level aggregate columns
	std::vector<TypeProps> final_aggrs;  // after re-aggregation (count -> sum)
};

struct Lowered {
	Typing ty;
	Plan plan;
	std::vector<BuildPlan> builds;
	std::vector<uint64_t> capacity;      // join table slots
	std::vector<uint64_t> dir_words;     // direct layout: 64-bit words of dir[] (0 = hash layout)
	std::vector<int> build_grid;
	std::vector<int> sum_of_aggr;        // aggregate -> sum index (-1: count)
	uint64_t key_range[VH_MAX_KEYS] = {};
	int key_bits[VH_MAX_KEYS] = {};
	uint64_t groups = 0;                 // output slots (domain or hash capacity)
	bool wide = false;
	int grid = 0;
	int block = 256;                     // generated probe kernel: threads per workgroup
	// radix-partitioned probe (Plan::radix): rows per lane of its probe
	// kernel, both launches' LDS bytes, the join kernel's workgroup size and
	// grid (one workgroup per partition)
	int radix_unroll = 0;
	size_t radix_lds_a = 0, radix_lds_b = 0;
	int radix_block = 1024;
	int radix_grid = 1;
};

struct Query {
	vh_pipeline_desc desc;
	vh_config cfg;
	Lowered L;
	u32* flags = nullptr;
	u64* slots[VH_MAX_JOINS] = {};
	u64* dir[VH_MAX_JOINS] = {};
	u64* ovf[VH_MAX_JOINS] = {};       // overflow table of a partitioned hash table
	u32* bits[VH_MAX_JOINS] = {};      // JoinT::lds == LDS_BITS: key bitmap of dir
	struct FillSeg* dclear = nullptr;   // segments of the per-run clear pass
	int nclear = 0;
	Plan* dplan = nullptr;
	BuildPlan* dbuild[VH_MAX_JOINS] = {};   // entries of dbuilds
	BuildPlan* dbuilds = nullptr;
	// two or more partitioned builds run as the work items of two launches
	// (scatter chunks, then slices), as a batch runs them
	struct BatchItem* dparts = nullptr;
	u32* dctl = nullptr;                     // the two launches' queue counters
	u32 nscatter = 0, nslice = 0;
	size_t slice_lds = 0;
	std::vector<void*> allocs;
	bool ran = false;
	std::string result;
	hipFunction_t jit_fn = nullptr;   // generated probe kernel (null: generic)
	hipEvent_t probe_start = nullptr;  // timestamps of the generated kernel's own start and end
	hipEvent_t probe_stop = nullptr;
	std::string jit_src;
	// radix-partitioned probe (Plan::radix): the join kernel after the probe
	// kernel and its source (one workgroup per partition)
	hipFunction_t jit_join_fn = nullptr;
	std::string jit_join_src;
	int generic_grid = 0;              // probe grid of the generic kernel
	bool fanout_rerun = false;         // switched to the generic kernel for a JoinN fan-out
	~Query();
	int alloc(void** p, size_t bytes);
};

// error word, then per join a 64-bit word: its repeated-key flag, then its
// spill flag (a partitioned hash build put rows in the overflow table)
static constexpr int kFlagWords = 2 + 2 * VH_MAX_JOINS;
inline u32* dups_flag(Query* q, int j) { return q->flags + 2 + 2 * j; }
inline u32* spill_flag(Query* q, int j) { return q->flags + 3 + 2 * j; }

int fail(int code, const char* fmt, ...);
const char* last_error();
int type_from_minmax(double mn, double mx);
int type_pipeline(const vh_pipeline_desc* d, Typing& ty);
int lower(const vh_pipeline_desc* d, const vh_config* cfg, Lowered& L);
// GB_PART: tuples of a region (Plan::gregion_rows) staged in LDS
int part_stage(const Plan& P);
// GB_PART: slice size (gshift) and partitions (gparts), at least `want`
// partitions where slices of 64 groups allow
void part_partitions(Plan& P, i64 want);
// single: a query of its own (generated probe kernel per cfg.kernel,
// partitioned builds); false for the pipelines of a batch
int query_create(Query* q, bool single = true);
// Shape of a generated probe kernel: the join tables it keeps in LDS
// (JoinT::lds, lds_off) and its workgroup size (L.block, L.grid)
void choose_probe_shape(Lowered& L, const vh_pipeline_desc* d, const vh_config* cfg);
int query_run(Query* q, hipStream_t s, int phase);
int query_result(Query* q, hipStream_t s, std::string& text, int64_t& nrows);

} // namespace vh
