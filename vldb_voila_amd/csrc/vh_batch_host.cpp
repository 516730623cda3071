// Host side of the batched launch: lowers every pipeline like a single
// query, lays out the work queue (all build items, then all probe items) and
// enqueues one multi-buffer clear plus persistent launches (builds, then the
// probe items of each generated kernel or generic group-by strategy, then
// partition aggregations), or replays them from a captured hipGraph.  Probe
// items run generated kernels (vh_generated_batch, one per distinct source:
// pipelines that differ only in data share one) unless a pipeline's config
// asks for the generic kernel or hipRTC is unavailable.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/voila_hip.h"
#include "vh_batch.hpp"
#include "vh_engine.hpp"
#include "vh_jit.hpp"
#include "vh_kernels.hpp"

using namespace vh;

#define HIP_CALL(x)                                                                  \
	do {                                                                             \
		hipError_t _e = (x);                                                         \
		if (_e != hipSuccess)                                                        \
			return fail(VH_ERR_HIP, "%s: %s", #x, hipGetErrorString(_e));           \
	} while (0)

struct BatchLaunch {
	int kind;            // kBatchBuild, kBatchAgg, kBatchGenerated or a GroupMode (probe items)
	u32 first;
	u32 count;
	size_t lds;
	hipFunction_t fn;    // kBatchGenerated: the compiled vh_generated_batch
};

static constexpr int kBatchGenerated = -3;
static constexpr int kBatchScatter = -4;   // partitioned builds: chunk items
static constexpr int kBatchSlice = -5;     // partitioned builds: partition items
static constexpr int kBatchPostFill = -6;  // the fills that wait for the builds' flags
// probe items' grouping key (BatchItem::build) of generated kernel g
static constexpr int kGenKey = 1000;

struct vh_batch {
	std::vector<Query*> q;
	std::vector<void*> allocs;
	Plan* d_plans = nullptr;
	BuildPlan* d_builds = nullptr;
	BatchItem* d_items = nullptr;
	FillSeg* d_segs = nullptr;
	FillSeg* d_psegs = nullptr;   // fills after the builds (conditional on their repeated-key flags)
	int npsegs = 0;
	u32* d_ctl = nullptr;
	u32 nitems = 0;
	std::vector<BatchLaunch> launches;
	size_t agg_lds = 0;
	int nsegs = 0;
	int grid = 0;
	size_t lds = 0;
	size_t mode_lds[8] = {};
	hipGraphExec_t graph = nullptr;
	hipGraph_t graph_def = nullptr;
	hipStream_t graph_stream = nullptr;
	~vh_batch()
	{
		if (graph)
			(void)hipGraphExecDestroy(graph);
		if (graph_def)
			(void)hipGraphDestroy(graph_def);
		for (Query* x : q)
			delete x;
		for (void* p : allocs)
			(void)hipFree(p);
	}
	int upload(void** dst, const void* src, size_t bytes)
	{
		hipError_t e = hipMalloc(dst, std::max<size_t>(bytes, 16));
		if (e != hipSuccess)
			return fail(VH_ERR_NOMEM, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
		allocs.push_back(*dst);
		if (src && bytes && (e = hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice)) != hipSuccess)
			return fail(VH_ERR_HIP, "hipMemcpy: %s", hipGetErrorString(e));
		return VH_OK;
	}
};

// Rows of a probe work item: 16 K by default, 8 K for a partitioned
// group-by, whose item stages its first 32 KiB of tuples in LDS and sends
// the rest through memory (measured on the 256-pipeline batch, staged
// tuples: 4.22 ms at 8 K rows, 4.26 at 4 K; 4.42 and 4.58 with 48 KiB
// staged, which leaves fewer workgroups per CU; without the stage: 4.88 at
// 16 K)
static int tile_rows_of(const vh_config& c, bool part = false)
{
	int shift = c.tile_shift;
	if (shift < -6) shift = -6;
	if (shift > 6) shift = 6;
	const int base = part ? 8192 : 16384;
	return shift >= 0 ? (base << shift) : (base >> (-shift));
}

// Lowering of one pipeline of a batch: lane-private group-by state cannot
// outlive a work item (LDS instead), and a partitioned group-by's regions
// are the work items.
static int lower_for_batch(int i, const vh_pipeline_desc* d, const vh_config& cfg, Lowered& L)
{
	vh_config lc = cfg;
	if (lc.group_mode == VH_GROUP_REG)
		lc.group_mode = VH_GROUP_LDS;
	lc.radix = -1;   // batched probe items run the one-pass kernels
	int rc = lower(d, &lc, L);
	if (rc)
		return fail(rc, "pipeline %d: %s", i, last_error());
	Plan& P = L.plan;
	if (P.mode == GB_REG)
		P.mode = GB_LDS;
	i64 T = tile_rows_of(cfg, P.mode == GB_PART);
	if (P.mode == GB_PART) {
		// the default item size grows for a pipeline that would otherwise
		// have more regions than a partition's run table holds
		while (cfg.tile_shift == 0 && (P.rows + T - 1) / T > kMaxGroupRegions && T < ((i64)1 << 30))
			T <<= 1;
		P.gregion_rows = T;
		P.gstage = part_stage(P);
		// one aggregation item per partition, each paying a fixed cost (its
		// runs' table, the slice's clear and store): as few partitions as
		// the slice size allows, more for a pipeline with many rows, about
		// 32 K rows each (batch: 4.22 ms at 32 K rows per partition, 4.32 at
		// 16 K, 4.48 at 64 K, 4.39 with ~256 partitions per pipeline)
		part_partitions(P, (P.rows + 32767) / 32768);
		if ((P.rows + T - 1) / T > kMaxGroupRegions)
			return fail(VH_ERR_UNSUPPORTED, "pipeline %d: more than %d items for a partitioned group-by", i,
				kMaxGroupRegions);
	}
	return VH_OK;
}

// clear, then every build item, then the probe items of each group-by
// strategy (a launch each), then the partition aggregation items of
// partitioned group-bys
static hipError_t enqueue(vh_batch* b, hipStream_t s)
{
	hipError_t e = launch_multi_fill(b->d_segs, b->nsegs, s);
	for (size_t k = 0; k < b->launches.size() && e == hipSuccess; k++) {
		const BatchLaunch& l = b->launches[k];
		if (l.kind == kBatchPostFill) {
			e = launch_multi_fill(b->d_psegs, b->npsegs, s);
		} else if (l.kind == kBatchScatter || l.kind == kBatchSlice) {
			e = launch_batch_part(l.kind == kBatchSlice, b->d_builds, b->d_items, l.first, l.count, b->d_ctl + k,
				std::max(1, b->grid / 4), l.lds, s);
		} else if (l.kind == kBatchGenerated) {
			if (!l.count)
				continue;
			const Plan* plans = b->d_plans;
			const BatchItem* items = b->d_items;
			u32 first = l.first, count = l.count;
			u32* counter = b->d_ctl + k;
			void* args[] = {(void*)&plans, (void*)&items, (void*)&first, (void*)&count, (void*)&counter};
			e = hipModuleLaunchKernel(l.fn, (unsigned)b->grid, 1, 1, 256, 1, 1, (unsigned)l.lds, s, args, nullptr);
		} else {
			e = launch_batch(l.kind, b->d_plans, b->d_builds, b->d_items, l.first, l.count, b->d_ctl + k, b->grid,
				l.lds, s);
		}
	}
	return e;
}

extern "C" {

int vh_batch_new(vh_batch** out, int32_t n, const vh_pipeline_desc* descs, const vh_config* cfgs, int32_t grid)
{
	if (!out || n <= 0 || !descs)
		return fail(VH_ERR_INVALID, "vh_batch_new arguments");
	*out = nullptr;
	vh_batch* b = new (std::nothrow) vh_batch();
	if (!b)
		return fail(VH_ERR_NOMEM, "host allocation");
	std::vector<Plan> plans;
	std::vector<BuildPlan> builds;
	std::vector<BatchItem> build_items, probe_items, agg_items, scatter_items, slice_items;
	size_t slice_lds = 0;
	std::vector<FillSeg> segs, psegs;
	std::vector<std::string> gen_src;        // generated batch kernels, by index
	std::vector<hipFunction_t> gen_fn;
	std::vector<size_t> gen_lds;
	int rc = VH_OK;
	for (int i = 0; i < n && !rc; i++) {
		Query* q = new (std::nothrow) Query();
		if (!q) { rc = fail(VH_ERR_NOMEM, "host allocation"); break; }
		b->q.push_back(q);
		memcpy(&q->desc, &descs[i], sizeof(vh_pipeline_desc));
		if (cfgs) {
			q->cfg = cfgs[i];
		} else {
			memset(&q->cfg, 0, sizeof(q->cfg));
			q->cfg.group_mode = VH_GROUP_AUTO;
		}
		if ((rc = lower_for_batch(i, &q->desc, q->cfg, q->L)))
			break;
		Plan& P = q->L.plan;
		// rows per work item (a partitioned group-by's regions, as lowered)
		const i64 T = P.mode == GB_PART ? P.gregion_rows : (i64)tile_rows_of(q->cfg);
		if ((rc = query_create(q, false))) {
			fail(rc, "pipeline %d: %s", i, last_error());
			break;
		}
		b->mode_lds[P.mode] = std::max(b->mode_lds[P.mode], pipeline_lds_bytes(P));
		const int first_build = (int)builds.size();
		for (int j = 0; j < P.njoins; j++) {
			const BuildPlan& B = q->L.builds[j];
			builds.push_back(B);
			if (B.nparts) {
				// a direct table built in two LDS passes, as a single query's
				for (int c = 0; c < B.nchunks; c++)
					scatter_items.push_back({i, 0, first_build + j, 0, c, c + 1});
				for (int p = 0; p < B.nparts; p++)
					slice_items.push_back({i, 0, first_build + j, 0, p, p + 1});
				slice_lds = std::max(slice_lds, part_slice_lds_bytes(B));
			} else {
				for (i64 r0 = 0; r0 < B.rows; r0 += T)
					build_items.push_back({i, 0, first_build + j, 0, r0, std::min<i64>(r0 + T, B.rows)});
			}
			// as a single query clears them: a partitioned hash table's
			// slices store every slot, so its overflow table is refilled
			// after a run that spilled into it; a partitioned direct build
			// rewrites all of dir[], and its overflow slots are refilled
			// after a run with repeated keys (the scatter launch resets both
			// flags); a CAS build's slots, dir[] and flag every run
			if (B.hashed) {
				segs.push_back({q->ovf[j], q->L.capacity[j], ~0ull, spill_flag(q, j)});
			} else if (B.nparts) {
				segs.push_back({q->slots[j], q->L.capacity[j], ~0ull, dups_flag(q, j)});
			} else {
				segs.push_back({(u64*)dups_flag(q, j), 1, 0, nullptr});
				segs.push_back({q->slots[j], q->L.capacity[j], ~0ull, nullptr});
				if (q->dir[j])
					segs.push_back({q->dir[j], q->L.dir_words[j], ~0ull, nullptr});
			}
		}
		// the probe kernel: generated for this pipeline's plan (shared with
		// every pipeline whose source is the same), or generic per strategy
		int key = P.mode;
		if (q->cfg.kernel != VH_KERNEL_GENERIC) {
			const std::string src = jit_source(P, true, jit_unroll(&q->cfg), true, 256, true);
			size_t g = 0;
			while (g < gen_src.size() && gen_src[g] != src)
				g++;
			if (g == gen_src.size()) {
				hipFunction_t fn = nullptr;
				std::string log;
				const int jrc = jit_compile(src, &fn, log, "vh_generated_batch");
				if (jrc && q->cfg.kernel != VH_KERNEL_AUTO) {
					rc = fail(jrc, "pipeline %d: %s", i, last_error());
					break;
				}
				if (!jrc) {
					gen_src.push_back(src);
					gen_fn.push_back(fn);
					gen_lds.push_back(0);
				}
			}
			if (g < gen_src.size()) {
				key = kGenKey + (int)g;
				gen_lds[g] = std::max(gen_lds[g], pipeline_lds_bytes(P));
			}
		}
		for (i64 r0 = 0; r0 < P.rows; r0 += T)
			probe_items.push_back({i, 1, key, 0, r0, std::min<i64>(r0 + T, P.rows)});
		if (P.mode == GB_PART) {
			for (int p = 0; p < P.gparts; p++)
				agg_items.push_back({i, 2, -1, 0, p, p + 1});
			const size_t R = (size_t)((P.rows + T - 1) / T);
			b->agg_lds = std::max(b->agg_lds, (part_slice_bytes(P) + 7) / 8 * 8 + (2 * R + 1) * sizeof(u32));
		}
		const u64 groups = q->L.groups;
		const u64 ns = (u64)std::max(P.nsums, 1);
		segs.push_back({(u64*)q->flags, 1, 0, nullptr});   // error bits
		if (P.mode == GB_PART) {
			// the partition aggregation stores every group's count (and the
			// sums of the groups present); the arrays need zeros only when a
			// JoinN fan-out can add rows into them during the probe, which
			// takes repeated keys in that JoinN's build of this run (the
			// segments of two such joins may both clear them)
			for (int j = 0; j < P.njoins; j++) {
				if (P.joins[j].variant != VH_JOINN)
					continue;
				const u32* f = dups_flag(q, j);
				psegs.push_back({P.out_lo, groups * ns, 0, f});
				psegs.push_back({(u64*)P.out_hi, groups * ns, 0, f});
				psegs.push_back({P.out_cnt, groups, 0, f});
			}
		} else {
			segs.push_back({P.out_lo, groups * ns, 0});
			segs.push_back({(u64*)P.out_hi, groups * ns, 0});
			segs.push_back({P.out_cnt, groups, 0});
			if (P.mode == GB_HASH)
				segs.push_back({P.hkeys, groups, ~0ull});
		}
		plans.push_back(P);
	}
	if (rc) {
		delete b;
		return rc;
	}
	// queue: build items, probe items grouped by group-by strategy (the
	// probe item's build field holds its plan's mode), aggregation items
	std::stable_sort(probe_items.begin(), probe_items.end(),
		[](const BatchItem& x, const BatchItem& y) { return x.build < y.build; });
	std::vector<BatchItem> items = build_items;
	b->launches.push_back({kBatchBuild, 0, (u32)build_items.size(), 0, nullptr});
	b->launches.push_back({kBatchScatter, (u32)items.size(), (u32)scatter_items.size(), 0, nullptr});
	items.insert(items.end(), scatter_items.begin(), scatter_items.end());
	b->launches.push_back({kBatchSlice, (u32)items.size(), (u32)slice_items.size(), slice_lds, nullptr});
	items.insert(items.end(), slice_items.begin(), slice_items.end());
	if (!psegs.empty())
		b->launches.push_back({kBatchPostFill, 0, 0, 0, nullptr});
	const size_t nbuild_items = items.size();
	for (size_t k = 0; k < probe_items.size();) {
		size_t e = k;
		while (e < probe_items.size() && probe_items[e].build == probe_items[k].build)
			e++;
		const int key = probe_items[k].build;
		if (key >= kGenKey)
			b->launches.push_back({kBatchGenerated, (u32)(nbuild_items + k), (u32)(e - k),
				gen_lds[key - kGenKey], gen_fn[key - kGenKey]});
		else
			b->launches.push_back({key, (u32)(nbuild_items + k), (u32)(e - k), b->mode_lds[key], nullptr});
		k = e;
	}
	items.insert(items.end(), probe_items.begin(), probe_items.end());
	// aggregation items, longest first (rows per partition of their
	// pipeline): the workgroups that take the last items then wait less
	// for the longest ones
	std::stable_sort(agg_items.begin(), agg_items.end(), [&](const BatchItem& x, const BatchItem& y) {
		const Plan& a = plans[x.plan];
		const Plan& c = plans[y.plan];
		return (double)a.rows / a.gparts > (double)c.rows / c.gparts;
	});
	b->launches.push_back({kBatchAgg, (u32)items.size(), (u32)agg_items.size(), b->agg_lds, nullptr});
	items.insert(items.end(), agg_items.begin(), agg_items.end());
	for (const BatchLaunch& l : b->launches)
		if (l.kind != kBatchBuild)
			b->lds = std::max(b->lds, l.lds);
	b->nitems = (u32)items.size();
	const size_t ctl_words = (b->launches.size() + 1) / 2 + 1;   // queue counters (u32), one per launch
	if ((rc = b->upload((void**)&b->d_ctl, nullptr, ctl_words * 8)) ||
		(rc = b->upload((void**)&b->d_plans, plans.data(), plans.size() * sizeof(Plan))) ||
		(rc = b->upload((void**)&b->d_builds, builds.data(), builds.size() * sizeof(BuildPlan))) ||
		(rc = b->upload((void**)&b->d_items, items.data(), items.size() * sizeof(BatchItem)))) {
		delete b;
		return rc;
	}
	segs.push_back({(u64*)b->d_ctl, ctl_words, 0});
	b->nsegs = (int)segs.size();
	b->npsegs = (int)psegs.size();
	if ((rc = b->upload((void**)&b->d_segs, segs.data(), segs.size() * sizeof(FillSeg))) ||
		(!psegs.empty() && (rc = b->upload((void**)&b->d_psegs, psegs.data(), psegs.size() * sizeof(FillSeg))))) {
		delete b;
		return rc;
	}
	int dev = 0, cus = 256;
	(void)hipGetDevice(&dev);
	(void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
	// eight persistent workgroups per CU (measured on the 256-pipeline batch:
	// 5.35 ms at 8 per CU, 5.53 at 4, 6.60 at 2, 5.47 at 24; a workgroup that
	// finds its queue drained returns at once)
	b->grid = grid > 0 ? grid : cus * 8;
	*out = b;
	return VH_OK;
}

int vh_generate_batch_source(const vh_pipeline_desc* desc, const vh_config* cfg, char* buf, size_t cap,
	size_t* len)
{
	if (!desc || !len)
		return fail(VH_ERR_INVALID, "vh_generate_batch_source arguments");
	vh_config c;
	if (cfg) {
		c = *cfg;
	} else {
		memset(&c, 0, sizeof(c));
		c.group_mode = VH_GROUP_AUTO;
	}
	Lowered L;
	int rc = lower_for_batch(0, desc, c, L);
	if (rc)
		return rc;
	const std::string src = jit_source(L.plan, true, jit_unroll(&c), true, 256, true);
	*len = src.size();
	if (buf) {
		if (cap <= src.size())
			return fail(VH_ERR_CAPACITY, "source needs %zu bytes", src.size() + 1);
		memcpy(buf, src.data(), src.size());
		buf[src.size()] = 0;
	}
	return VH_OK;
}

int vh_batch_run(vh_batch* b, void* stream)
{
	if (!b)
		return fail(VH_ERR_INVALID, "null batch");
	hipStream_t s = (hipStream_t)stream;
	HIP_CALL(enqueue(b, s));
	for (Query* q : b->q)
		q->ran = true;
	return VH_OK;
}

// Capture vh_batch_run once into a hipGraph (on `stream`, which must not be
// the null stream) and replay it on every call.
int vh_batch_run_graph(vh_batch* b, void* stream)
{
	if (!b || !stream)
		return fail(VH_ERR_INVALID, "vh_batch_run_graph needs a batch and a non-null stream");
	hipStream_t s = (hipStream_t)stream;
	if (!b->graph) {
		HIP_CALL(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
		hipError_t e1 = enqueue(b, s);
		hipGraph_t g = nullptr;
		HIP_CALL(hipStreamEndCapture(s, &g));
		if (e1 != hipSuccess) {
			(void)hipGraphDestroy(g);
			return fail(VH_ERR_HIP, "capture of the batch failed");
		}
		b->graph_def = g;
		HIP_CALL(hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0));
	}
	HIP_CALL(hipGraphLaunch(b->graph, s));
	for (Query* q : b->q)
		q->ran = true;
	return VH_OK;
}

int vh_batch_result(vh_batch* b, int32_t i, void* stream, char* buf, size_t cap, size_t* len, int64_t* nrows)
{
	if (!b || i < 0 || i >= (int)b->q.size() || !len)
		return fail(VH_ERR_INVALID, "vh_batch_result arguments");
	HIP_CALL(hipStreamSynchronize((hipStream_t)stream));
	int64_t rows = 0;
	Query* q = b->q[i];
	int rc = query_result(q, (hipStream_t)stream, q->result, rows);
	if (rc)
		return rc;
	*len = q->result.size();
	if (nrows)
		*nrows = rows;
	if (buf) {
		if (cap <= q->result.size())
			return fail(VH_ERR_CAPACITY, "result needs %zu bytes", q->result.size() + 1);
		memcpy(buf, q->result.data(), q->result.size());
		buf[q->result.size()] = 0;
	}
	return VH_OK;
}

int vh_batch_info(const vh_batch* b, int32_t* items, int32_t* grid, int64_t* lds_bytes)
{
	if (!b)
		return fail(VH_ERR_INVALID, "null batch");
	if (items) *items = (int32_t)b->nitems;
	if (grid) *grid = b->grid;
	if (lds_bytes) *lds_bytes = (int64_t)b->lds;
	return VH_OK;
}

int vh_batch_free(vh_batch* b)
{
	delete b;
	return VH_OK;
}

} // extern "C"
