// Fused gfx950 kernels for scan -> select -> hash join* -> hash group-by.
//
// One pass over the probe columns evaluates the selection predicates, walks
// the open-addressing join tables built by build_kernel, computes the sum
// arguments in registers and aggregates them:
//   GB_REG     dense group index, lane-private LDS accumulators (<= 8 groups)
//   GB_LDS     dense group index into a per-block LDS table (LDS atomics)
//   GB_GLOBAL  dense group index straight into global atomics
//   GB_HASH    packed group key into a global CAS open-addressing table
//   GB_LDSHASH per-block LDS hash table over the dense index, global overflow
// GB_REG / GB_LDS blocks publish exact per-block partials; the last block to
// finish merges them (no atomics on the result, no clearing pass).
//
// The statement sequences this replaces are VOILA's select (vec_<cmp> +
// vec_seltrue), the probe loop (bucket_lookup / check / gather / bucket_next,
// relalg_translator.cpp:933-1111) and the group-by loop (hash, bucket_lookup,
// check, bucket_insert, scatter, aggr_sum / aggr_count,
// relalg_translator.cpp:582-656).
#include "vh_device.hpp"
#include "vh_fused.hpp"
#include "vh_plan.hpp"

namespace vh {

static constexpr int kBlock = 256;

// Exact merge of the per-block partials, one workgroup per output element.
// Partials are element-major: sum element e (= s * domain + g) keeps its low
// words at partials[(2e) * nb + b] and high words at partials[(2e+1) * nb + b],
// count g at partials[(2 * domain * nsums + g) * nb + b], b = block.  A launch
// of its own after the pipeline kernel: no block waits for another, so no
// device-scope fence or completion counter sits on the hot path.
__global__ __launch_bounds__(kBlock) void merge_kernel(const Plan* __restrict__ Pp, u32 nb)
{
	const Plan& P = *Pp;
	__shared__ u64 red_lo[kBlock / kWave];
	__shared__ i64 red_hi[kBlock / kWave];
	const u64 D = P.domain;
	const u64 Dn = D * (u64)P.nsums;
	const u64 e = blockIdx.x;
	const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
	u64 lo = 0;
	i64 hi = 0;
	if (e < Dn) {
		const u64* pl = P.partials + 2 * e * nb;
		const u64* ph = pl + nb;
		for (u32 b = threadIdx.x; b < nb; b += blockDim.x) {
			const u64 l = pl[b];
			const u64 nl = lo + l;
			hi += (i64)ph[b] + (nl < lo ? 1 : 0);
			lo = nl;
		}
	} else {
		const u64* pc = P.partials + (2 * Dn + (e - Dn)) * nb;
		for (u32 b = threadIdx.x; b < nb; b += blockDim.x)
			lo += pc[b];
	}
	wave_sum_i128(lo, hi);
	if (lane == 0) {
		red_lo[wave] = lo;
		red_hi[wave] = hi;
	}
	__syncthreads();
	if (threadIdx.x != 0)
		return;
	lo = 0;
	hi = 0;
	for (int w = 0; w < (int)(blockDim.x / kWave); w++) {
		const u64 nl = lo + red_lo[w];
		hi += red_hi[w] + (nl < lo ? 1 : 0);
		lo = nl;
	}
	if (e < Dn) {
		P.out_lo[e] = lo;
		P.out_hi[e] = hi;
	} else {
		P.out_cnt[e - Dn] = lo;
	}
}

template <int MODE>
__global__ __launch_bounds__(kBlock) void pipeline_kernel(const Plan* __restrict__ Pp)
{
	const Plan& P = *Pp;
	extern __shared__ u64 lds[];

	const Lds L = lds_layout<MODE>(P, lds);
	lds_init<MODE>(P, L);
	u32 dups_mask = 0;
	for (int j = 0; j < P.njoins; j++)
		if (*P.joins[j].dups)
			dups_mask |= 1u << j;
	__syncthreads();

	const i64 tile = (i64)kU * blockDim.x;
	process_range<MODE>(P, L, (i64)blockIdx.x * tile, P.rows, (i64)gridDim.x * tile, dups_mask);

	block_epilogue<MODE>(P, L);
}

__global__ __launch_bounds__(kBlock) void build_kernel(const BuildPlan* __restrict__ Bp)
{
	const BuildPlan& B = *Bp;
	const i64 tile = (i64)kU * blockDim.x;
	for (i64 base = (i64)blockIdx.x * tile; base < B.rows; base += (i64)gridDim.x * tile)
		build_chunk(B, base, B.rows);
}

__global__ __launch_bounds__(kBlock) void fill_u64_kernel(u64* p, u64 n, u64 v, const u32* only_if)
{
	if (only_if && *only_if == 0u)
		return;
	const u64 stride = (u64)gridDim.x * blockDim.x;
	for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
		p[i] = v;
}

// ---- launchers ----------------------------------------------------------

hipError_t launch_fill(u64* p, u64 n, u64 v, hipStream_t s, const u32* only_if)
{
	if (!n)
		return hipSuccess;
	u64 blocks = (n + kBlock - 1) / kBlock;
	if (blocks > 4096)
		blocks = 4096;
	hipLaunchKernelGGL(fill_u64_kernel, dim3((unsigned)blocks), dim3(kBlock), 0, s, p, n, v, only_if);
	return hipGetLastError();
}

hipError_t launch_build(const BuildPlan* B, int grid, hipStream_t s)
{
	hipLaunchKernelGGL(build_kernel, dim3(grid), dim3(kBlock), 0, s, B);
	return hipGetLastError();
}

size_t pipeline_lds_bytes(const Plan& P)
{
	if (P.mode == GB_REG)
		return (size_t)(P.domain * P.nsums * kBlock * 12 + P.domain * kBlock * 4 + 16);
	if (P.mode == GB_LDS)
		return (size_t)(P.domain * P.nsums * 12 + P.domain * 4 + 16);
	if (P.mode == GB_LDSHASH)
		return (size_t)P.nsums * kLdsHashSlots * 12 + kLdsHashSlots * 8 + 16;
	return 0;
}

hipError_t launch_merge(const Plan& P, const Plan* dP, int grid, hipStream_t s)
{
	if (P.mode != GB_REG && P.mode != GB_LDS)
		return hipSuccess;
	const u64 elems = P.domain * (u64)P.nsums + P.domain;
	hipLaunchKernelGGL(merge_kernel, dim3((unsigned)elems), dim3(kBlock), 0, s, dP, (u32)grid);
	return hipGetLastError();
}

hipError_t launch_pipeline(const Plan& P, const Plan* dP, int grid, hipStream_t s)
{
	const size_t lds = pipeline_lds_bytes(P);
#define VH_LAUNCH(M) hipLaunchKernelGGL((pipeline_kernel<M>), dim3(grid), dim3(kBlock), lds, s, dP)
	switch (P.mode) {
	case GB_REG: VH_LAUNCH(GB_REG); break;
	case GB_LDS: VH_LAUNCH(GB_LDS); break;
	case GB_GLOBAL: VH_LAUNCH(GB_GLOBAL); break;
	case GB_LDSHASH: VH_LAUNCH(GB_LDSHASH); break;
	default: VH_LAUNCH(GB_HASH); break;
	}
#undef VH_LAUNCH
	hipError_t e = hipGetLastError();
	if (e != hipSuccess)
		return e;
	return launch_merge(P, dP, grid, s);
}

} // namespace vh
