// Fused gfx950 kernels for scan -> select -> hash join* -> hash group-by.
//
// One pass over the probe columns evaluates the selection predicates, walks
// the open-addressing join tables built by build_kernel, computes the sum
// arguments in registers and aggregates them:
//   GB_REG    dense group index into per-lane registers (<= 8 groups),
//   GB_LDS    dense group index into a per-block LDS table,
//   GB_GLOBAL dense group index straight into global atomics,
//   GB_HASH   packed group key into a global CAS open-addressing table.
// GB_REG / GB_LDS blocks publish exact per-block partials; the last block to
// finish merges them (no atomics on the result, no clearing pass).
//
// The statement sequences this replaces are VOILA's select (vec_<cmp> +
// vec_seltrue), the probe loop (bucket_lookup / check / gather / bucket_next,
// relalg_translator.cpp:933-1111) and the group-by loop (hash, bucket_lookup,
// check, bucket_insert, scatter, aggr_sum / aggr_count,
// relalg_translator.cpp:582-656).
#include "vh_device.hpp"
#include "vh_plan.hpp"

namespace vh {

static constexpr u64 kEmpty = ~0ull;

struct PayRegs {
	i64 p[VH_MAX_JOINS];
};

__device__ __forceinline__ i64 role_value(const Role& r, i64 row, const PayRegs& pay)
{
	switch (r.kind) {
	case ROLE_COL:
		return load1(r.ptr, r.type, row);
	case ROLE_PAY:
		switch (r.slot) {
		case 0: return pay.p[0];
		case 1: return pay.p[1];
		case 2: return pay.p[2];
		default: return pay.p[3];
		}
	case ROLE_CONST:
		return r.c;
	default:
		return 0;
	}
}

__device__ __forceinline__ u64 eval_sum(const SumArg& s, i64 row, const PayRegs& pay)
{
	u64 prod = 1;
#pragma unroll
	for (int t = 0; t < 3; t++) {
		if (t < s.nterms) {
			const Term& T = s.t[t];
			i64 v = T.c;
			if (T.sa)
				v += T.sa > 0 ? role_value(T.a, row, pay) : -role_value(T.a, row, pay);
			if (T.sb)
				v += T.sb > 0 ? role_value(T.b, row, pay) : -role_value(T.b, row, pay);
			prod = t == 0 ? (u64)v : prod * (u64)v;
		}
	}
	return prod;
}

__device__ __forceinline__ bool preds_pass(const Pred* preds, int npreds, i64 row)
{
	PayRegs none = {};
	bool ok = true;
#pragma unroll
	for (int p = 0; p < VH_MAX_PREDS; p++) {
		if (p < npreds) {
			const Pred& P = preds[p];
			ok = ok && compare(P.cmp, P.r.type, role_value(P.r, row, none), P.value);
		}
	}
	return ok;
}

// First match of a join whose build keys are unique (Join01, or JoinN whose
// build reported no duplicate key).
__device__ __forceinline__ bool probe_first(const JoinT& J, i64 key, i64& pay)
{
	if (key < J.kmin || key > J.kmax)
		return false;
	const u32 off = (u32)(key - J.kmin);
	u64 h = voila_hash((u64)off) & J.mask;
	for (u64 n = 0; n <= J.mask; n++) {
		const u64 s = J.slots[h];
		if (s == kEmpty)
			return false;
		if ((u32)s == off) {
			pay = (i64)(s >> 32) + J.pmin;
			return true;
		}
		h = (h + 1) & J.mask;
	}
	return false;
}

// Per-lane state of the aggregation.
template <int MODE, bool WIDE>
struct Acc {
	u64 lo[kRegSlots][kRegSums];
	int hi[kRegSlots][kRegSums];
	u32 cnt[kRegSlots];
};

template <int MODE, bool WIDE>
__device__ __forceinline__ void consume(const Plan& P, i64 row, const PayRegs& pay,
	Acc<MODE, WIDE>& A, u64* lds_lo, int* lds_hi, u32* lds_cnt, u32* lds_keys)
{
	if (MODE == GB_HASH) {
		u64 packed = 0;
		bool in_range = true;
#pragma unroll
		for (int k = 0; k < VH_MAX_KEYS; k++) {
			if (k < P.nkeys) {
				const u64 off = (u64)(role_value(P.keys[k].r, row, pay) - P.keys[k].min);
				in_range = in_range && off <= P.keys[k].range;
				packed |= off << P.keys[k].shift;
			}
		}
		if (!in_range) {
			atomicOr(P.error, 8u);   // key outside its column's [min, max]
			return;
		}
		u64 h = voila_hash(packed) & P.hmask;
		u64 n = 0;
		for (; n <= P.hmask; n++) {
			u64 k = P.hkeys[h];
			if (k == packed)
				break;
			if (k == kEmpty) {
				u64 old = atomicCAS(&P.hkeys[h], kEmpty, packed);
				if (old == kEmpty || old == packed)
					break;
			}
			h = (h + 1) & P.hmask;
		}
		if (n > P.hmask) {
			atomicOr(P.error, 1u);
			return;
		}
		const u64 groups = P.hmask + 1;
		atomicAdd(&P.out_cnt[h], 1ull);
#pragma unroll
		for (int s = 0; s < VH_MAX_AGGRS; s++) {
			if (s < P.nsums) {
				u64 v = eval_sum(P.sums[s], row, pay);
				if (WIDE && P.sums[s].res128)
					atomic_add_i128(&P.out_lo[s * groups + h], &P.out_hi[s * groups + h], v,
						(P.sums[s].arg_signed && (i64)v < 0) ? -1 : 0);
				else
					atomicAdd(&P.out_lo[s * groups + h], v);
			}
		}
		return;
	}

	u64 g = 0;
	bool in_range = true;
#pragma unroll
	for (int k = 0; k < VH_MAX_KEYS; k++) {
		if (k < P.nkeys) {
			const u64 off = (u64)(role_value(P.keys[k].r, row, pay) - P.keys[k].min);
			in_range = in_range && off <= P.keys[k].range;
			g += off * P.keys[k].mult;
		}
	}
	if (!in_range || g >= P.domain) {
		atomicOr(P.error, 8u);   // key outside its column's [min, max]
		return;
	}

	if (MODE == GB_REG) {
		u64 v[kRegSums];
#pragma unroll
		for (int s = 0; s < kRegSums; s++)
			v[s] = s < P.nsums ? eval_sum(P.sums[s], row, pay) : 0;
#pragma unroll
		for (int d = 0; d < kRegSlots; d++) {
			if ((u64)d < P.domain) {
				const bool m = g == (u64)d;
				A.cnt[d] += m ? 1u : 0u;
#pragma unroll
				for (int s = 0; s < kRegSums; s++) {
					if (s < P.nsums) {
						const u64 add = m ? v[s] : 0;
						const u64 o = A.lo[d][s];
						A.lo[d][s] = o + add;
						if (WIDE && P.sums[s].wide)
							A.hi[d][s] += (int)((o + add) < o) -
								(int)(P.sums[s].arg_signed && (i64)add < 0);
					}
				}
			}
		}
		return;
	}

	if (MODE == GB_LDS) {
		atomicAdd(&lds_cnt[g], 1u);
#pragma unroll
		for (int s = 0; s < VH_MAX_AGGRS; s++) {
			if (s < P.nsums) {
				const u64 v = eval_sum(P.sums[s], row, pay);
				const u64 idx = (u64)s * P.domain + g;
				if (WIDE && P.sums[s].wide) {
					const u64 o = atomicAdd(&lds_lo[idx], v);
					const int h = (int)((o + v) < o) - (int)(P.sums[s].arg_signed && (i64)v < 0);
					if (h)
						atomicAdd(&lds_hi[idx], h);
				} else {
					atomicAdd(&lds_lo[idx], v);
				}
			}
		}
		return;
	}

	if (MODE == GB_LDSHASH) {
		const u32 key = (u32)g;
		u32 slot = (u32)(voila_hash((u64)key)) & (kLdsHashSlots - 1);
		int found = -1;
		for (int n = 0; n < kLdsHashProbes; n++) {
			const u32 k = lds_keys[slot];
			if (k == key) { found = (int)slot; break; }
			if (k == 0xFFFFFFFFu) {
				const u32 old = atomicCAS(&lds_keys[slot], 0xFFFFFFFFu, key);
				if (old == 0xFFFFFFFFu || old == key) { found = (int)slot; break; }
			}
			slot = (slot + 1) & (kLdsHashSlots - 1);
		}
		if (found >= 0) {
			atomicAdd(&lds_cnt[found], 1u);
#pragma unroll
			for (int s = 0; s < VH_MAX_AGGRS; s++) {
				if (s < P.nsums) {
					const u64 v = eval_sum(P.sums[s], row, pay);
					const int idx = s * kLdsHashSlots + found;
					if (WIDE && P.sums[s].wide) {
						const u64 o = atomicAdd(&lds_lo[idx], v);
						const int h = (int)((o + v) < o) - (int)(P.sums[s].arg_signed && (i64)v < 0);
						if (h)
							atomicAdd(&lds_hi[idx], h);
					} else {
						atomicAdd(&lds_lo[idx], v);
					}
				}
			}
			return;
		}
		// LDS table full: fall through to the global arrays
	}

	// GB_GLOBAL
	atomicAdd(&P.out_cnt[g], 1ull);
#pragma unroll
	for (int s = 0; s < VH_MAX_AGGRS; s++) {
		if (s < P.nsums) {
			const u64 v = eval_sum(P.sums[s], row, pay);
			const u64 idx = (u64)s * P.domain + g;
			if (WIDE && P.sums[s].res128)
				atomic_add_i128(&P.out_lo[idx], &P.out_hi[idx], v,
					(P.sums[s].arg_signed && (i64)v < 0) ? -1 : 0);
			else
				atomicAdd(&P.out_lo[idx], v);
		}
	}
}

template <int MODE, bool WIDE>
__device__ __forceinline__ void process_row(const Plan& P, i64 row, Acc<MODE, WIDE>& A,
	u64* lds_lo, int* lds_hi, u32* lds_cnt, u32* lds_keys)
{
	if (!preds_pass(P.preds, P.npreds, row))
		return;
	PayRegs pay = {};
	const int last = P.njoins - 1;
#pragma unroll
	for (int j = 0; j < VH_MAX_JOINS - 1; j++) {
		if (j < last) {
			const JoinT& J = P.joins[j];
			i64 p = 0;
			if (!probe_first(J, role_value(J.key, row, pay), p))
				return;
			pay.p[j] = p;
		}
	}
	if (P.njoins == 0) {
		consume<MODE, WIDE>(P, row, pay, A, lds_lo, lds_hi, lds_cnt, lds_keys);
		return;
	}
	// last join: every match continues (JoinN) or the first one (Join01)
	const JoinT& J = P.joins[last];
	const i64 key = role_value(J.key, row, pay);
	if (key < J.kmin || key > J.kmax)
		return;
	const u32 off = (u32)(key - J.kmin);
	u64 h = voila_hash((u64)off) & J.mask;
	for (u64 n = 0; n <= J.mask; n++) {
		const u64 s = J.slots[h];
		if (s == kEmpty)
			return;
		if ((u32)s == off) {
			const i64 p = (i64)(s >> 32) + J.pmin;
			switch (last) {
			case 0: pay.p[0] = p; break;
			case 1: pay.p[1] = p; break;
			case 2: pay.p[2] = p; break;
			default: pay.p[3] = p; break;
			}
			consume<MODE, WIDE>(P, row, pay, A, lds_lo, lds_hi, lds_cnt, lds_keys);
			if (J.variant == VH_JOIN01)
				return;
		}
		h = (h + 1) & J.mask;
	}
}

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
		u64 lo = 0;
		i64 hi = 0;
		if (e < E && (int)threadIdx.x / G < (int)blockDim.x / G) {
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
		if (sub == 0 && e < E && (int)threadIdx.x / G < (int)blockDim.x / G) {
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

template <int MODE, bool WIDE>
__global__ __launch_bounds__(256) void pipeline_kernel(const Plan* __restrict__ Pp)
{
	const Plan& P = *Pp;
	extern __shared__ u64 lds[];
	__shared__ u64 red_lo[256];
	__shared__ i64 red_hi[256];
	__shared__ u32 is_last;

	Acc<MODE, WIDE> A;
	u64* lds_lo = nullptr;
	int* lds_hi = nullptr;
	u32* lds_cnt = nullptr;
	u32* lds_keys = nullptr;

	if (MODE == GB_REG) {
#pragma unroll
		for (int d = 0; d < kRegSlots; d++) {
			A.cnt[d] = 0;
#pragma unroll
			for (int s = 0; s < kRegSums; s++) {
				A.lo[d][s] = 0;
				A.hi[d][s] = 0;
			}
		}
	}
	if (MODE == GB_LDS) {
		const u64 n = P.domain * P.nsums;
		lds_lo = lds;
		lds_hi = (int*)(lds + n);
		lds_cnt = (u32*)(lds_hi + n);
		for (u64 i = threadIdx.x; i < n; i += blockDim.x) {
			lds_lo[i] = 0;
			lds_hi[i] = 0;
		}
		for (u64 i = threadIdx.x; i < P.domain; i += blockDim.x)
			lds_cnt[i] = 0;
		__syncthreads();
	}

	if (MODE == GB_LDSHASH) {
		const int n = P.nsums * kLdsHashSlots;
		lds_lo = lds;
		lds_hi = (int*)(lds + n);
		lds_cnt = (u32*)(lds_hi + n);
		lds_keys = lds_cnt + kLdsHashSlots;
		for (int i = threadIdx.x; i < n; i += blockDim.x) {
			lds_lo[i] = 0;
			lds_hi[i] = 0;
		}
		for (int i = threadIdx.x; i < kLdsHashSlots; i += blockDim.x) {
			lds_cnt[i] = 0;
			lds_keys[i] = 0xFFFFFFFFu;
		}
		__syncthreads();
	}

	const i64 stride = (i64)gridDim.x * blockDim.x;
	constexpr int U = 4;
	for (i64 base = (i64)blockIdx.x * blockDim.x + threadIdx.x; base < P.rows; base += stride * U) {
#pragma unroll
		for (int u = 0; u < U; u++) {
			const i64 row = base + u * stride;
			if (row < P.rows)
				process_row<MODE, WIDE>(P, row, A, lds_lo, lds_hi, lds_cnt, lds_keys);
		}
	}

	if (MODE == GB_LDSHASH) {
		// flush this block's LDS groups into the global arrays
		__syncthreads();
		for (int i = threadIdx.x; i < kLdsHashSlots; i += blockDim.x) {
			const u32 key = lds_keys[i];
			if (key == 0xFFFFFFFFu)
				continue;
			atomicAdd(&P.out_cnt[key], (u64)lds_cnt[i]);
			for (int s = 0; s < P.nsums; s++) {
				const SumArg& S = P.sums[s];
				const u64 lo = lds_lo[s * kLdsHashSlots + i];
				const i64 hi = (WIDE && S.wide) ? (i64)lds_hi[s * kLdsHashSlots + i]
					: (S.arg_signed ? (i64)lo >> 63 : 0);
				const u64 idx = (u64)s * P.domain + key;
				if (S.res128)
					atomic_add_i128(&P.out_lo[idx], &P.out_hi[idx], lo, hi);
				else
					atomicAdd(&P.out_lo[idx], lo);
			}
		}
		return;
	}

	if (MODE != GB_REG && MODE != GB_LDS)
		return;

	// ---- publish this block's exact partials --------------------------------
	const u64 D = P.domain;
	const u64 W = 2 * D * P.nsums + D;
	u64* mine = P.partials + (u64)blockIdx.x * W;
	if (MODE == GB_REG) {
		const int lane = threadIdx.x % kWave;
		const int wave = threadIdx.x / kWave;
		const int nwaves = blockDim.x / kWave;
		// wave sums into LDS [wave][entry], entry = s * D + d or counts
		u64* wl = lds;                                      // nwaves * E words (lo)
		i64* wh = (i64*)(lds + nwaves * kRegSlots * (kRegSums + 1));
#pragma unroll
		for (int d = 0; d < kRegSlots; d++) {
			if ((u64)d < D) {
#pragma unroll
				for (int s = 0; s < kRegSums; s++) {
					if (s < P.nsums) {
						u64 lo = A.lo[d][s];
						// a narrow lane sum fits in 64 bits: sign extend it
						i64 hi = (WIDE && P.sums[s].wide) ? (i64)A.hi[d][s]
							: (P.sums[s].arg_signed ? (i64)lo >> 63 : 0);
						wave_sum_i128(lo, hi);
						if (lane == 0) {
							wl[wave * kRegSlots * (kRegSums + 1) + s * kRegSlots + d] = lo;
							wh[wave * kRegSlots * (kRegSums + 1) + s * kRegSlots + d] = hi;
						}
					}
				}
				const u64 c = wave_sum_u64(A.cnt[d]);
				if (lane == 0)
					wl[wave * kRegSlots * (kRegSums + 1) + kRegSums * kRegSlots + d] = c;
			}
		}
		__syncthreads();
		const int E = (int)(D * P.nsums + D);
		for (int e = threadIdx.x; e < E; e += blockDim.x) {
			u64 lo = 0;
			i64 hi = 0;
			int idx;
			if ((u64)e < D * P.nsums)
				idx = (int)(e / D) * kRegSlots + (int)(e % D);
			else
				idx = kRegSums * kRegSlots + (int)((u64)e - D * P.nsums);
			for (int w = 0; w < nwaves; w++) {
				const u64 l = wl[w * kRegSlots * (kRegSums + 1) + idx];
				const u64 nl = lo + l;
				if ((u64)e < D * P.nsums)
					hi += wh[w * kRegSlots * (kRegSums + 1) + idx] + (nl < lo ? 1 : 0);
				lo = nl;
			}
			if ((u64)e < D * P.nsums) {
				mine[2 * e] = lo;
				mine[2 * e + 1] = (u64)hi;
			} else {
				mine[2 * D * P.nsums + ((u64)e - D * P.nsums)] = lo;
			}
		}
	} else {
		__syncthreads();
		const u64 n = D * P.nsums;
		for (u64 i = threadIdx.x; i < n; i += blockDim.x) {
			const SumArg& S = P.sums[i / D];
			const u64 lo = lds_lo[i];
			mine[2 * i] = lo;
			mine[2 * i + 1] = (WIDE && S.wide) ? (u64)(i64)lds_hi[i]
				: (S.arg_signed ? (u64)((i64)lo >> 63) : 0ull);
		}
		for (u64 i = threadIdx.x; i < D; i += blockDim.x)
			mine[2 * n + i] = lds_cnt[i];
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

// Build side: select, then insert (key - kmin, payload - pmin) with CAS into
// linear-probing slots.  Replaces write_pos/write + bucket_build
// (relalg_translator.cpp:828-928, runtime_struct.cpp:1283-1324).
__global__ __launch_bounds__(256) void build_kernel(const BuildPlan* __restrict__ Bp)
{
	const BuildPlan& B = *Bp;
	const i64 stride = (i64)gridDim.x * blockDim.x;
	for (i64 row = (i64)blockIdx.x * blockDim.x + threadIdx.x; row < B.rows; row += stride) {
		if (!preds_pass(B.preds, B.npreds, row))
			continue;
		const i64 key = load1(B.key, B.key_type, row);
		if (key < B.kmin || key > B.kmax) {
			atomicOr(B.error, 4u);
			continue;
		}
		const u32 off = (u32)(key - B.kmin);
		const u32 pay = B.npay ? (u32)(load1(B.pay, B.pay_type, row) - B.pmin) : 0u;
		const u64 val = (u64)off | ((u64)pay << 32);
		u64 h = voila_hash((u64)off) & B.mask;
		for (u64 n = 0; n <= B.mask; n++) {
			const u64 old = atomicCAS(&B.slots[h], kEmpty, val);
			if (old == kEmpty)
				break;
			if ((u32)old == off)
				*B.dups = 1u;
			h = (h + 1) & B.mask;
		}
	}
}

__global__ __launch_bounds__(256) void fill_u64_kernel(u64* p, u64 n, u64 v)
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
	u64 blocks = (n + 255) / 256;
	if (blocks > 4096)
		blocks = 4096;
	hipLaunchKernelGGL(fill_u64_kernel, dim3((unsigned)blocks), dim3(256), 0, s, p, n, v);
	return hipGetLastError();
}

hipError_t launch_build(const BuildPlan* B, int grid, hipStream_t s)
{
	hipLaunchKernelGGL(build_kernel, dim3(grid), dim3(256), 0, s, B);
	return hipGetLastError();
}

size_t pipeline_lds_bytes(const Plan& P)
{
	if (P.mode == GB_LDSHASH)
		return (size_t)P.nsums * kLdsHashSlots * 12 + kLdsHashSlots * 8 + 16;
	if (P.mode == GB_LDS)
		return P.domain * P.nsums * 12 + P.domain * 4 + 16;
	if (P.mode == GB_REG)
		return (size_t)(256 / kWave) * kRegSlots * (kRegSums + 1) * 16;
	return 0;
}

hipError_t launch_pipeline(const Plan& P, const Plan* dP, int grid, bool wide, hipStream_t s)
{
	const size_t lds = pipeline_lds_bytes(P);
#define VH_LAUNCH(M, W) hipLaunchKernelGGL((pipeline_kernel<M, W>), dim3(grid), dim3(256), lds, s, dP)
	switch (P.mode) {
	case GB_REG: if (wide) VH_LAUNCH(GB_REG, true); else VH_LAUNCH(GB_REG, false); break;
	case GB_LDS: if (wide) VH_LAUNCH(GB_LDS, true); else VH_LAUNCH(GB_LDS, false); break;
	case GB_GLOBAL: if (wide) VH_LAUNCH(GB_GLOBAL, true); else VH_LAUNCH(GB_GLOBAL, false); break;
	case GB_LDSHASH: if (wide) VH_LAUNCH(GB_LDSHASH, true); else VH_LAUNCH(GB_LDSHASH, false); break;
	default: if (wide) VH_LAUNCH(GB_HASH, true); else VH_LAUNCH(GB_HASH, false); break;
	}
#undef VH_LAUNCH
	return hipGetLastError();
}

} // namespace vh
