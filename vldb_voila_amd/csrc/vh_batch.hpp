// Work items and launchers of the persistent batched kernel (vh_batch.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "vh_device.hpp"
#include "vh_plan.hpp"

namespace vh {

// p[0, n) = v; with only_if, only when *only_if != 0 when the launch reads it
struct FillSeg {
	u64* p;
	u64 n;
	u64 v;
	const u32* only_if;
};

// One launch filling every segment (the batch's and a query's clear pass).
hipError_t launch_multi_fill(const FillSeg* d_segs, int nseg, hipStream_t s);
// Launch kinds: build items, partition-aggregation items, or probe items of
// one group-by strategy (the GroupMode value).
static constexpr int kBatchBuild = -1;
static constexpr int kBatchAgg = -2;
// Items [first, first + count) of one kind with the launch's own queue counter.
hipError_t launch_batch(int kind, const Plan* d_plans, const BuildPlan* d_builds, const BatchItem* d_items,
	u32 first, u32 count, u32* d_counter, int grid, size_t lds_bytes, hipStream_t s);

// A workgroup's next item of a persistent launch, or false when its queue is
// empty (*counter is the launch's cursor, *slot a word of the workgroup's LDS).
// For launches of large, uneven items; the build launches, whose items are
// small, hold two at a time (queue_begin in vh_fused.hpp).
__device__ __forceinline__ bool next_batch_item(const BatchItem* __restrict__ items, u32 first, u32 count,
	u32* __restrict__ counter, u32* slot, BatchItem& I)
{
	if (threadIdx.x == 0)
		*slot = gatomic_add(counter, 1u);
	__syncthreads();
	const u32 k = *slot;
	__syncthreads();
	if (k >= count)
		return false;
	I = items[first + k];
	return true;
}

} // namespace vh
