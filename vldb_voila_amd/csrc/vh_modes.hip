// The generic fused kernels of one group-by strategy.  This file is compiled
// once per strategy with -DVH_MODE=<GroupMode>, so the six instantiations of
// the plan-interpreting pipeline kernel (and the five of the batched probe
// kernel) build in parallel translation units; vh_pipeline.hip and
// vh_batch.hip dispatch to them by strategy.
#include "vh_batch.hpp"
#include "vh_device.hpp"
#include "vh_fused.hpp"
#include "vh_kernels.hpp"
#include "vh_plan.hpp"

#ifndef VH_MODE
#error "vh_modes.hip is compiled with -DVH_MODE=<GroupMode>"
#endif

namespace vh {

static constexpr int kModeBlock = 256;

template <int MODE>
__global__ __launch_bounds__(kModeBlock) void pipeline_kernel(const Plan* __restrict__ Pp)
{
	const Plan& P = *Pp;
	extern __shared__ u64 lds[];

	Lds L = lds_layout<MODE>(P, lds);
	lds_init<MODE>(P, L);
	const u32 dups_mask = join_flags(P);
	__syncthreads();

	const i64 tile = (i64)kU * blockDim.x;
	if (MODE == GB_PART) {
		// a contiguous slab of rows per workgroup: its region of tuple space
		const i64 r0 = (i64)blockIdx.x * P.gregion_rows;
		const i64 r1 = r0 + P.gregion_rows < P.rows ? r0 + P.gregion_rows : P.rows;
		L.rbase = r0;
		L.rcap = r1 > r0 ? (u32)(r1 - r0) : 0u;
		L.region = blockIdx.x;
		process_range<MODE>(P, L, r0, r1, tile, dups_mask);
	} else {
		process_range<MODE>(P, L, (i64)blockIdx.x * tile, P.rows, (i64)gridDim.x * tile, dups_mask);
	}

	block_epilogue<MODE>(P, L);
}

template <>
hipError_t launch_pipeline_mode<VH_MODE>(const Plan* dP, int grid, size_t lds, hipStream_t s)
{
	hipLaunchKernelGGL((pipeline_kernel<VH_MODE>), dim3(grid), dim3(kModeBlock), lds, s, dP);
	return hipGetLastError();
}

#if VH_MODE != GB_REG
template <int MODE>
__device__ __forceinline__ void probe_item_m(const Plan& P, i64 r0, i64 r1, u64* lds, u32 dups_mask)
{
	Lds L = lds_layout<MODE>(P, lds);
	if (MODE == GB_PART) {
		// an item is one region: its rows' tuple space and table row
		L.rbase = r0;
		L.rcap = (u32)(r1 - r0);
		L.region = (u32)(r0 / P.gregion_rows);
	}
	lds_init<MODE>(P, L);
	__syncthreads();
	process_range<MODE>(P, L, r0, r1, (i64)kU * blockDim.x, dups_mask);
	__syncthreads();
	if (MODE == GB_LDS)
		lds_dense_flush(P, L);
	else if (MODE == GB_LDSHASH)
		ldshash_flush(P, L);
	else if (MODE == GB_PART)
		part_finish(P, L);
	__syncthreads();
}

template <int MODE>
__global__ __launch_bounds__(kModeBlock) void batch_probe_kernel(const Plan* __restrict__ plans,
	const BatchItem* __restrict__ items, u32 first, u32 count, u32* __restrict__ counter)
{
	extern __shared__ u64 lds[];
	__shared__ u32 slot;
	BatchItem I;
	while (next_batch_item(items, first, count, counter, &slot, I)) {
		const Plan& P = plans[I.plan];
		const u32 dups_mask = join_flags(P);
		probe_item_m<MODE>(P, I.r0, I.r1, lds, dups_mask);
	}
}

template <>
hipError_t launch_batch_probe_mode<VH_MODE>(const Plan* d_plans, const BatchItem* d_items, u32 first, u32 count,
	u32* d_counter, int grid, size_t lds_bytes, hipStream_t s)
{
	hipLaunchKernelGGL((batch_probe_kernel<VH_MODE>), dim3(grid), dim3(kModeBlock), lds_bytes, s, d_plans, d_items,
		first, count, d_counter);
	return hipGetLastError();
}
#endif

} // namespace vh
