// Persistent batched execution of many independent pipelines.
//
// The reference runs pipelines strictly one after the other
// (runtime_framework.cpp:172-198), each with its own flavor
// (QueryConfig::pipeline_default_blend, runtime_framework.hpp:78).  Here the
// build and probe work of every pipeline is cut into row ranges ("items").
// One persistent launch works through all build items, a second one through
// all probe items; each workgroup takes the next item from an atomic counter
// until its queue is empty.  No workgroup ever waits for another: the second
// launch starts after the first has finished, which orders every build
// before every probe.  Group-by state lives in LDS per item and is added to
// the pipeline's global result arrays when the item ends.  Both launches and
// the clearing pass are captured into one hipGraph by vh_batch_run_graph.
#include "vh_batch.hpp"
#include "vh_device.hpp"
#include "vh_fused.hpp"
#include "vh_kernels.hpp"
#include "vh_plan.hpp"

namespace vh {

// Items [first, first + count) of the queue; *counter is this launch's
// cursor.  One kernel per kind of item (and per group-by strategy), so each
// launch holds only the registers its own code needs.
__global__ __launch_bounds__(256) void batch_build_kernel(const BuildPlan* __restrict__ builds,
	const BatchItem* __restrict__ items, u32 first, u32 count, u32* __restrict__ counter)
{
	__shared__ u32 qs[2];
	u32 cur, nxt;
	queue_begin(counter, qs, cur, nxt);
	BatchItem I = items[first + (cur < count ? cur : 0)];
	while (cur < count) {
		const u32 req = queue_request(counter);
		const BatchItem In = items[first + (nxt < count ? nxt : cur)];
		const BuildPlan& B = builds[I.build];
		for (i64 base = I.r0; base < I.r1; base += (i64)kU * blockDim.x)
			build_chunk(B, base, I.r1);
		const u32 nn = queue_advance(qs, req);
		cur = nxt;
		nxt = nn;
		I = In;
	}
}

// GB_PART partition aggregation: item r0 = partition; LDS = the slice, then
// two per-region arrays and a total.
__global__ __launch_bounds__(256) void batch_agg_kernel(const Plan* __restrict__ plans,
	const BatchItem* __restrict__ items, u32 first, u32 count, u32* __restrict__ counter)
{
	extern __shared__ u64 lds[];
	__shared__ u32 slot;
	BatchItem I;
	while (next_batch_item(items, first, count, counter, &slot, I)) {
		const Plan& P = plans[I.plan];
		const u64 slice_words = (((u64)1 << P.gshift) * ((u64)P.nsums * 12 + 4) + 7) / 8;
		const int R = (int)((P.rows + P.gregion_rows - 1) / P.gregion_rows);
		u32* pre = (u32*)(lds + slice_words);
		part_agg<4>(P, (int)I.r0, lds, pre, pre + R, pre + 2 * R);
		__syncthreads();
	}
}

// One launch that fills many buffers (join slots, result arrays): keeps a
// captured graph at two nodes however many pipelines the batch holds.
__global__ __launch_bounds__(256) void multi_fill_kernel(const FillSeg* __restrict__ segs, int nseg)
{
	for (int s = blockIdx.y; s < nseg; s += gridDim.y) {
		const FillSeg S = segs[s];
		if (S.only_if && *S.only_if == 0u)
			continue;
		for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < S.n; i += (u64)gridDim.x * blockDim.x)
			S.p[i] = S.v;
	}
}

hipError_t launch_multi_fill(const FillSeg* d_segs, int nseg, hipStream_t s)
{
	if (nseg <= 0)
		return hipSuccess;
	const unsigned gy = nseg < 65535 ? (unsigned)nseg : 65535u;
	hipLaunchKernelGGL(multi_fill_kernel, dim3(64, gy), dim3(256), 0, s, d_segs, nseg);
	return hipGetLastError();
}

hipError_t launch_batch(int kind, const Plan* d_plans, const BuildPlan* d_builds, const BatchItem* d_items,
	u32 first, u32 count, u32* d_counter, int grid, size_t lds_bytes, hipStream_t s)
{
	if (!count)
		return hipSuccess;
#define VH_PROBE(M) return launch_batch_probe_mode<M>(d_plans, d_items, first, count, d_counter, grid, lds_bytes, s)
	switch (kind) {
	case kBatchBuild:
		hipLaunchKernelGGL(batch_build_kernel, dim3(grid), dim3(256), 0, s, d_builds, d_items, first, count,
			d_counter);
		break;
	case kBatchAgg:
		hipLaunchKernelGGL(batch_agg_kernel, dim3(grid), dim3(256), lds_bytes, s, d_plans, d_items, first, count,
			d_counter);
		break;
	case GB_LDS: VH_PROBE(GB_LDS);
	case GB_LDSHASH: VH_PROBE(GB_LDSHASH);
	case GB_GLOBAL: VH_PROBE(GB_GLOBAL);
	case GB_PART: VH_PROBE(GB_PART);
	default: VH_PROBE(GB_HASH);
	}
#undef VH_PROBE
	return hipGetLastError();
}

} // namespace vh
