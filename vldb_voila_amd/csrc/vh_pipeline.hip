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

// Exact merge of per-block partials by the last block to finish.
// Partial layout per block: lo/hi of (s * domain + g) for all sums, then cnt[g].
__device__ void merge_partials(const Plan& P, u64* red_lo, i64* red_hi)
{
	const u64 D = P.domain;
	const int E = (int)(D * P.nsums + D);
	const u64 W = 2 * D * P.nsums + D;        // words per block partial
	const int nb = gridDim.x;
	int G = (int)blockDim.x / E;
	if (G < 1)
		G = 1;
	for (int base = 0; base < E; base += (int)blockDim.x / G) {
		const int e = base + (int)threadIdx.x / G;
		const int sub = (int)threadIdx.x % G;
		const bool mine = e < E && (int)threadIdx.x / G < (int)blockDim.x / G;
		u64 lo = 0;
		i64 hi = 0;
		if (mine) {
			if ((u64)e < D * P.nsums) {
				for (int b = sub; b < nb; b += G) {
					const u64* pb = P.partials + (u64)b * W;
					const u64 l = __atomic_load_n(&pb[2 * e], __ATOMIC_RELAXED);
					const i64 h = (i64)__atomic_load_n(&pb[2 * e + 1], __ATOMIC_RELAXED);
					const u64 nl = lo + l;
					hi += h + (nl < lo ? 1 : 0);
					lo = nl;
				}
			} else {
				const u64 g = (u64)e - D * P.nsums;
				for (int b = sub; b < nb; b += G)
					lo += __atomic_load_n(&P.partials[(u64)b * W + 2 * D * P.nsums + g], __ATOMIC_RELAXED);
			}
		}
		red_lo[threadIdx.x] = lo;
		red_hi[threadIdx.x] = hi;
		__syncthreads();
		if (sub == 0 && mine) {
			for (int k = 1; k < G; k++) {
				const u64 l = red_lo[threadIdx.x + k];
				const u64 nl = lo + l;
				hi += red_hi[threadIdx.x + k] + (nl < lo ? 1 : 0);
				lo = nl;
			}
			if ((u64)e < D * P.nsums) {
				P.out_lo[e] = lo;
				P.out_hi[e] = hi;
			} else {
				P.out_cnt[(u64)e - D * P.nsums] = lo;
			}
		}
		__syncthreads();
	}
}

template <int MODE>
__global__ __launch_bounds__(kBlock) void pipeline_kernel(const Plan* __restrict__ Pp)
{
	const Plan& P = *Pp;
	extern __shared__ u64 lds[];
	__shared__ u64 red_lo[kBlock];
	__shared__ i64 red_hi[kBlock];
	__shared__ u32 is_last;

	const Lds L = lds_layout<MODE>(P, lds);
	lds_init<MODE>(P, L);
	u32 dups_mask = 0;
	for (int j = 0; j < P.njoins; j++)
		if (*P.joins[j].dups)
			dups_mask |= 1u << j;
	__syncthreads();

	const i64 tile = (i64)kU * blockDim.x;
	process_range<MODE>(P, L, (i64)blockIdx.x * tile, P.rows, (i64)gridDim.x * tile, dups_mask);

	if (MODE == GB_LDSHASH) {
		__syncthreads();
		ldshash_flush(P, L);
		return;
	}
	if (MODE != GB_REG && MODE != GB_LDS)
		return;

	// ---- publish this block's exact partials --------------------------------
	__syncthreads();
	const u64 D = P.domain;
	const u64 W = 2 * D * P.nsums + D;
	u64* mine = P.partials + (u64)blockIdx.x * W;
	if (MODE == GB_REG) {
		const int lane = threadIdx.x % kWave;
		const int wave = threadIdx.x / kWave;
		const int nwaves = blockDim.x / kWave;
		const int E = (int)(D * P.nsums + D);
		const u32 bd = blockDim.x;
		for (int e = wave; e < E; e += nwaves) {
			if ((u64)e < D * P.nsums) {
				const u64 s = (u64)e / D, g = (u64)e % D;
				const SumArg& S = P.sums[s];
				u64 lo = 0;
				i64 hi = 0;
				for (u32 t = lane; t < bd; t += kWave) {
					const u64 idx = (g * P.nsums + s) * bd + t;
					const u64 l = L.lo[idx];
					// a lane's sum of <= 32-bit arguments fits 64 bits: sign extend it
					const i64 h = (P.wide && S.wide) ? (i64)L.hi[idx] : (S.arg_signed ? (i64)l >> 63 : 0);
					const u64 nl = lo + l;
					hi += h + (nl < lo ? 1 : 0);
					lo = nl;
				}
				wave_sum_i128(lo, hi);
				if (lane == 0) {
					mine[2 * e] = lo;
					mine[2 * e + 1] = (u64)hi;
				}
			} else {
				const u64 g = (u64)e - D * P.nsums;
				u64 c = 0;
				for (u32 t = lane; t < bd; t += kWave)
					c += L.cnt[g * bd + t];
				c = wave_sum_u64(c);
				if (lane == 0)
					mine[2 * D * P.nsums + g] = c;
			}
		}
	} else {
		const u64 n = D * P.nsums;
		for (u64 i = threadIdx.x; i < n; i += blockDim.x) {
			const SumArg& S = P.sums[i / D];
			const u64 lo = L.lo[i];
			mine[2 * i] = lo;
			mine[2 * i + 1] = (P.wide && S.wide) ? (u64)(i64)L.hi[i]
				: (S.arg_signed ? (u64)((i64)lo >> 63) : 0ull);
		}
		for (u64 i = threadIdx.x; i < D; i += blockDim.x)
			mine[2 * n + i] = L.cnt[i];
	}

	// ---- last block merges --------------------------------------------------
	__builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
	__syncthreads();
	if (threadIdx.x == 0) {
		const u32 prev = __hip_atomic_fetch_add(P.done, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
		is_last = prev == gridDim.x - 1;
	}
	__syncthreads();
	if (!is_last)
		return;
	__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
	merge_partials(P, red_lo, red_hi);
	if (threadIdx.x == 0)
		__hip_atomic_store(P.done, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ __launch_bounds__(kBlock) void build_kernel(const BuildPlan* __restrict__ Bp)
{
	const BuildPlan& B = *Bp;
	const i64 tile = (i64)kU * blockDim.x;
	for (i64 base = (i64)blockIdx.x * tile; base < B.rows; base += (i64)gridDim.x * tile)
		build_chunk(B, base, B.rows);
}

__global__ __launch_bounds__(kBlock) void fill_u64_kernel(u64* p, u64 n, u64 v)
{
	const u64 stride = (u64)gridDim.x * blockDim.x;
	for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
		p[i] = v;
}

// ---- launchers ----------------------------------------------------------

hipError_t launch_fill(u64* p, u64 n, u64 v, hipStream_t s)
{
	if (!n)
		return hipSuccess;
	u64 blocks = (n + kBlock - 1) / kBlock;
	if (blocks > 4096)
		blocks = 4096;
	hipLaunchKernelGGL(fill_u64_kernel, dim3((unsigned)blocks), dim3(kBlock), 0, s, p, n, v);
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
	return hipGetLastError();
}

} // namespace vh
