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
#include <algorithm>

#include "vh_device.hpp"
#include "vh_fused.hpp"
#include "vh_kernels.hpp"
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

static constexpr int kPartBlockAgg = 1024;   // workgroup size of part_agg_kernel

__global__ __launch_bounds__(kPartBlockAgg) void part_agg_kernel(const Plan* __restrict__ Pp)
{
	extern __shared__ u64 slice[];
	__shared__ u32 pre[kMaxGroupRegions];
	__shared__ u32 base[kMaxGroupRegions];
	__shared__ u32 total;
	part_agg<4>(*Pp, blockIdx.x, slice, pre, base, &total);
}

__global__ __launch_bounds__(kBlock) void build_kernel(const BuildPlan* __restrict__ Bp)
{
	const BuildPlan& B = *Bp;
	const i64 tile = (i64)kU * blockDim.x;
	for (i64 base = (i64)blockIdx.x * tile; base < B.rows; base += (i64)gridDim.x * tile)
		build_chunk(B, base, B.rows);
}

// ---- partitioned build of a direct table ---------------------------------
//
// Device-scope atomics are the cost of a direct table built in place: every
// row claims its entry with one, and they retire at a fixed device-wide rate
// however small the table.  The single-query path instead partitions the
// build rows by key range so that each partition's slice of the table fits
// in LDS, in two launches and without global atomics:
//   part_scatter_kernel  one workgroup per chunk of kPartChunk rows sorts
//                        its rows by partition in LDS (counting sort), writes
//                        them back contiguously and records, per partition,
//                        the run's length and offset in the chunk;
//   part_build_kernel    one workgroup per partition gathers its runs from
//                        every chunk, claims entries in an LDS copy of its
//                        slice (the claim rule of build_chunk, with LDS
//                        atomics) and stores the finished slice.
// Repeated keys still go to the overflow slots with global atomics.

static constexpr int kMaxParts = 4096;
static constexpr int kPartBlock = 1024;               // workgroup size of both launches
static constexpr int kPartChunk = kPartBlock * kU;    // rows per scatter workgroup
static constexpr int kMaxChunks = 2048;
static_assert(kMaxParts <= kWave * 64 && kMaxChunks <= kWave * 64, "block_exclusive_scan bound");
static_assert(kPartChunk == kBuildChunkRows && kMaxChunks == kMaxBuildChunks && kMaxParts == kMaxBuildParts,
	"partitioned build limits shared with the host");

// Linear-probing insert of a repeated key into the overflow slots.
__device__ __forceinline__ void overflow_insert(const BuildPlan& B, u32 off, u64 val)
{
	u64 h = voila_hash((u64)off) & B.mask;
	for (u64 n = 0;; n++) {
		if (n >= kMaxProbe || n > B.mask) {
			gatomic_or(B.error, 32u);
			return;
		}
		if (gatomic_cas(&B.slots[h], kEmpty, val) == kEmpty)
			return;
		h = (h + 1) & B.mask;
	}
}

// Chunk `chunk` of a partitioned build (the workgroup's LDS: hist, staged,
// total): select its rows, count them per partition, store them sorted by
// partition and record each partition's run.
__device__ __forceinline__ void part_scatter_chunk(const BuildPlan& B, int chunk, u32* hist, u64* staged,
	u32* total_slot)
{
	const int P = B.nparts;
	// the repeated-key flag of this build: set (if at all) by the slice
	// builds, which run after every chunk of this launch
	if (chunk == 0 && threadIdx.x == 0) {
		*B.dups = 0u;
		if (B.hashed)
			*B.spill = 0u;
	}
	for (int p = threadIdx.x; p < P; p += blockDim.x)
		hist[p] = 0;
	__syncthreads();
	const i64 r0 = (i64)chunk * kPartChunk;
	u32 off[kU], pay[kU], rank[kU];
	bool ok[kU];
	// one key and at most one payload column: every column of the kU rows is
	// requested before any is tested (rows past the end read row 0), so the
	// chunk costs one memory round trip; packed keys and payloads are read
	// by build_row after the predicates
	const bool packed = B.nkeys > 1 || B.npay > 1;
	i64 key[kU], pv[kU];
	if (!packed && B.key_type == VH_I32 && (!B.npay || B.pay_type == VH_I32)) {
		// int32 key and payload (or none): the 2 kU loads in straight-line
		// code, issued together (the typed loads below sit in a switch per
		// column and row, and each waits for its value)
		const int32_t* ps = B.npay ? (const int32_t*)B.pay : (const int32_t*)B.key;
		int32_t kk[kU], pp[kU];
#pragma unroll
		for (int u = 0; u < kU; u++) {
			const i64 r = r0 + (i64)u * blockDim.x + threadIdx.x;
			const i64 rr = r < B.rows ? r : 0;
			kk[u] = gload((const int32_t*)B.key, rr);
			pp[u] = gload(ps, rr);
			ok[u] = r < B.rows;
		}
#pragma unroll
		for (int u = 0; u < kU; u++) {
			key[u] = kk[u];
			pv[u] = B.npay ? (i64)pp[u] : B.pmin;
		}
	} else {
		i64 rr[kU];
#pragma unroll
		for (int u = 0; u < kU; u++) {
			const i64 r = r0 + (i64)u * blockDim.x + threadIdx.x;
			rr[u] = r < B.rows ? r : 0;
			ok[u] = r < B.rows;
			key[u] = B.kmin;
			pv[u] = B.pmin;
		}
		if (!packed) {
			loadn<kU>(B.key, B.key_type, rr, key);
			if (B.npay)
				loadn<kU>(B.pay, B.pay_type, rr, pv);
		}
	}
#pragma unroll
	for (int p = 0; p < VH_MAX_PREDS; p++) {
		if (p < B.npreds) {
			const Pred& Q = B.preds[p];
			i64 x[kU], rr[kU];
#pragma unroll
			for (int u = 0; u < kU; u++) {
				const i64 r = r0 + (i64)u * blockDim.x + threadIdx.x;
				rr[u] = r < B.rows ? r : 0;
			}
			loadn<kU>(Q.r.ptr, Q.r.type, rr, x);
#pragma unroll
			for (int u = 0; u < kU; u++)
				ok[u] = ok[u] && compare(Q.cmp, Q.r.type, x[u], Q.value);
		}
	}
#pragma unroll
	for (int u = 0; u < kU; u++) {
		if (packed) {
			build_row(B, r0 + (i64)u * blockDim.x + threadIdx.x, ok[u], off[u], pay[u]);
			continue;
		}
		if (ok[u] && (key[u] < B.kmin || key[u] > B.kmax)) {
			gatomic_or(B.error, 4u);
			ok[u] = false;
		}
		off[u] = (u32)(key[u] - B.kmin);
		pay[u] = (u32)(pv[u] - B.pmin);
	}
	// partition of each row: its key range (direct table) or the block of
	// 2^pshift slots its hash falls in (partitioned hash table)
	u32 part[kU];
#pragma unroll
	for (int u = 0; u < kU; u++)
		part[u] = B.hashed ? (u32)((voila_hash((u64)off[u]) & B.mask) >> B.pshift) : off[u] >> B.pshift;
#pragma unroll
	for (int u = 0; u < kU; u++)
		rank[u] = ok[u] ? atomicAdd(&hist[part[u]], 1u) : 0u;
	__syncthreads();
	const u64 G = (u64)B.nchunks;
	for (int p = threadIdx.x; p < P; p += blockDim.x)
		((VH_GLOBAL u32*)B.pcnt)[(u64)p * G + chunk] = hist[p];
	__syncthreads();
	const u32 n = block_exclusive_scan(hist, P, total_slot);
	for (int p = threadIdx.x; p < P; p += blockDim.x)
		((VH_GLOBAL u32*)B.poff)[(u64)p * G + chunk] = hist[p];
#pragma unroll
	for (int u = 0; u < kU; u++)
		if (ok[u])
			staged[hist[part[u]] + rank[u]] = (u64)off[u] | ((u64)pay[u] << 32);
	__syncthreads();
	VH_GLOBAL u64* dst = (VH_GLOBAL u64*)B.pbuf + r0;
	for (u32 i = threadIdx.x; i < n; i += blockDim.x)
		dst[i] = staged[i];
	__syncthreads();   // the LDS is free for the workgroup's next chunk
}

__global__ __launch_bounds__(kPartBlock) void part_scatter_kernel(const BuildPlan* __restrict__ Bp)
{
	__shared__ u32 hist[kMaxParts];
	__shared__ u64 staged[kPartChunk];
	__shared__ u32 total;
	part_scatter_chunk(*Bp, blockIdx.x, hist, staged, &total);
}

// Linear-probing insert into the overflow table of a partitioned hash table
// (a row whose sub-table is full); a slot holding the same key marks the
// build as having repeated keys.
__device__ __forceinline__ void spill_insert(const BuildPlan& B, u32 off, u64 val)
{
	*B.spill = 1u;
	u64 h = voila_hash((u64)off) & B.mask;
	for (u64 n = 0;; n++) {
		if (n >= kMaxProbe || n > B.mask) {
			gatomic_or(B.error, 32u);
			return;
		}
		const u64 old = gatomic_cas(&B.ovf[h], kEmpty, val);
		if (old == kEmpty)
			return;
		if ((u32)old == off)
			*B.dups = 1u;
		h = (h + 1) & B.mask;
	}
}

// Partition `p` of a partitioned hash table: its rows go into an LDS copy of
// its block of S = 2^pshift slots with LDS compare-and-swap and linear
// probing that wraps within the block (the probe kernels follow the same
// sequence); then the whole block is stored, so the table needs no clearing
// between runs.  A row that finds the block full goes to the overflow table.
__device__ __forceinline__ void part_build_hash_slice(const BuildPlan& B, int p, u64* table, u32* pre, u32* base,
	u32* total_rows)
{
	constexpr int V = 2 * kU;
	const int G = B.nchunks;
	const u32 S = 1u << B.pshift;
	for (u32 i = threadIdx.x; i < S; i += blockDim.x)
		table[i] = kEmpty;
	for (int b = threadIdx.x; b < G; b += blockDim.x) {
		pre[b] = gload(B.pcnt, (i64)p * G + b);
		base[b] = gload(B.poff, (i64)p * G + b);
	}
	__syncthreads();
	const u32 total = block_exclusive_scan(pre, G, total_rows);
	for (int b = threadIdx.x; b < G; b += blockDim.x)
		base[b] = (u32)b * kPartChunk + base[b] - pre[b];
	__syncthreads();
	bool repeated = false;
	for (u32 i0 = 0; i0 < total; i0 += V * blockDim.x) {
		u64 v[V];
		u32 iv[V];
		int rv[V];
#pragma unroll
		for (int u = 0; u < V; u++)
			iv[u] = i0 + u * blockDim.x + threadIdx.x;
		runs_of<V>(pre, G, iv, rv);
		u32 at[V];
#pragma unroll
		for (int u = 0; u < V; u++)
			at[u] = base[rv[u]];
#pragma unroll
		for (int u = 0; u < V; u++)
			at[u] = iv[u] < total ? at[u] + iv[u] : 0u;
#pragma unroll
		for (int u = 0; u < V; u++)
			v[u] = gload(B.pbuf, (i64)at[u]);
#pragma unroll
		for (int u = 0; u < V; u++)
			v[u] = iv[u] < total ? v[u] : kEmpty;
#pragma unroll
		for (int u = 0; u < V; u++) {
			if (v[u] == kEmpty)
				continue;
			const u32 off = (u32)v[u];
			u32 h = (u32)voila_hash((u64)off) & (S - 1);
			bool placed = false;
			for (u32 n = 0; n < S; n++) {
				const u64 old = atomicCAS(&table[h], kEmpty, v[u]);
				if (old == kEmpty) {
					placed = true;
					break;
				}
				if ((u32)old == off)
					repeated = true;
				h = (h + 1) & (S - 1);
			}
			if (!placed)
				spill_insert(B, off, v[u]);
		}
	}
	if (repeated)
		*B.dups = 1u;
	__syncthreads();
	VH_GLOBAL u64* dst = (VH_GLOBAL u64*)B.slots + (u64)p * S;
	for (u32 i = threadIdx.x; i < S; i += blockDim.x)
		dst[i] = table[i];
	__syncthreads();   // the LDS is free for the workgroup's next partition
}

// Partition `p` of a partitioned build: gather its runs from every chunk,
// claim entries in an LDS copy of its slice, store the slice (and its words
// of the key bitmap).  LDS: the slice, pre and base (kMaxChunks each), total.
__device__ __forceinline__ void part_build_slice(const BuildPlan& B, int p, u32* slice, u32* pre, u32* base,
	u32* total_rows)
{
	constexpr int V = 2 * kU;   // rows gathered per thread per step
	const int P = B.nparts;
	const int G = B.nchunks;
	const int w = B.dir_w;
	const u64 k0 = (u64)p << B.pshift;
	const u64 range = (u64)(B.kmax - B.kmin) + 1;
	const u64 nkeys = ((1ull << B.pshift) < range - k0) ? (1ull << B.pshift) : range - k0;
	const u32 words = (u32)((nkeys * (u64)w + 3) / 4);
	for (u32 i = threadIdx.x; i < words; i += blockDim.x)
		slice[i] = 0xFFFFFFFFu;
	for (int b = threadIdx.x; b < G; b += blockDim.x) {
		pre[b] = gload(B.pcnt, (i64)p * G + b);
		base[b] = gload(B.poff, (i64)p * G + b);
	}
	__syncthreads();
	const u32 total = block_exclusive_scan(pre, G, total_rows);
	// row i of this partition (0 <= i < total) is pbuf[base[b] + i] for the
	// last chunk b with pre[b] <= i
	for (int b = threadIdx.x; b < G; b += blockDim.x)
		base[b] = (u32)b * kPartChunk + base[b] - pre[b];
	__syncthreads();
	const u32 idle_key = (u32)(threadIdx.x % nkeys);
	for (u32 i0 = 0; i0 < total; i0 += V * blockDim.x) {
		u64 v[V];
		u32 iv[V];
		int rv[V];
#pragma unroll
		for (int u = 0; u < V; u++)
			iv[u] = i0 + u * blockDim.x + threadIdx.x;
		runs_of<V>(pre, G, iv, rv);
		// every lane loads (rows past the total read pbuf[0]), so the loads
		// are issued together
		u32 at[V];
#pragma unroll
		for (int u = 0; u < V; u++)
			at[u] = base[rv[u]];
#pragma unroll
		for (int u = 0; u < V; u++)
			at[u] = iv[u] < total ? at[u] + iv[u] : 0u;
#pragma unroll
		for (int u = 0; u < V; u++)
			v[u] = gload(B.pbuf, (i64)at[u]);
#pragma unroll
		for (int u = 0; u < V; u++)
			v[u] = iv[u] < total ? v[u] : kEmpty;
		// every lane's claims are issued before any result is tested (a row
		// past the total ANDs all-ones into a word of the slice chosen by its
		// thread, which changes nothing and keeps such no-ops off one word)
		u32* word[V];
		u32 mask[V], seen[V];
#pragma unroll
		for (int u = 0; u < V; u++) {
			const bool has = v[u] != kEmpty;
			u32 shift;
			word[u] = dir_word(slice, w, has ? (u32)v[u] - (u32)k0 : idle_key, mask[u], shift);
			const u32 bits = has ? ((((u32)(v[u] >> 32) << shift) & mask[u]) | ~mask[u]) : 0xFFFFFFFFu;
			seen[u] = atomicAnd(word[u], bits);
		}
#pragma unroll
		for (int u = 0; u < V; u++) {
			if (v[u] != kEmpty && (seen[u] & mask[u]) != mask[u]) {
				atomicOr(word[u], seen[u] & mask[u]);
				*B.dups = 1u;
				overflow_insert(B, (u32)v[u], v[u]);
			}
		}
	}
	__syncthreads();
	VH_GLOBAL u32* dst = (VH_GLOBAL u32*)((char*)B.dir + k0 * (u64)w);
	for (u32 i = threadIdx.x; i < words; i += blockDim.x)
		dst[i] = slice[i];
	if (B.bits) {
		// the slice's words of the key bitmap (k0 is a multiple of 2^10)
		const u32 empty = dir_empty(w);
		VH_GLOBAL u32* bdst = (VH_GLOBAL u32*)B.bits + (k0 >> 5);
		for (u32 i = threadIdx.x; i < (u32)((nkeys + 31) / 32); i += blockDim.x) {
			u32 word = 0;
			for (u32 b = 0; b < 32; b++) {
				const u64 k = (u64)i * 32 + b;
				if (k < nkeys && dir_load_lds(slice, w, (u32)k) != empty)
					word |= 1u << b;
			}
			bdst[i] = word;
		}
	}
	__syncthreads();   // the LDS is free for the workgroup's next partition
}

__device__ __forceinline__ void part_slice(const BuildPlan& B, int p, u64* lds_words, u32* pre, u32* base,
	u32* total_rows)
{
	if (B.hashed)
		part_build_hash_slice(B, p, lds_words, pre, base, total_rows);
	else
		part_build_slice(B, p, (u32*)lds_words, pre, base, total_rows);
}

__global__ __launch_bounds__(kPartBlock) void part_build_kernel(const BuildPlan* __restrict__ Bp)
{
	extern __shared__ u64 lds_words[];
	__shared__ u32 pre[kMaxChunks];
	__shared__ u32 base[kMaxChunks];
	__shared__ u32 total_rows;
	part_slice(*Bp, blockIdx.x, lds_words, pre, base, &total_rows);
}

// Batched partitioned builds: scatter items (build, chunk r0) and slice
// items (build, partition r0) from a queue, as the one-query launches do
// them by workgroup index; the slice launch runs after the scatter launch.
__global__ __launch_bounds__(kPartBlock) void batch_part_scatter_kernel(const BuildPlan* __restrict__ builds,
	const BatchItem* __restrict__ items, u32 first, u32 count, u32* __restrict__ counter)
{
	__shared__ u32 hist[kMaxParts];
	__shared__ u64 staged[kPartChunk];
	__shared__ u32 total;
	__shared__ u32 qs[2];
	u32 cur, nxt;
	queue_begin(counter, qs, cur, nxt);
	BatchItem I = items[first + (cur < count ? cur : 0)];
	while (cur < count) {
		const u32 req = queue_request(counter);
		const BatchItem In = items[first + (nxt < count ? nxt : cur)];
		part_scatter_chunk(builds[I.build], (int)I.r0, hist, staged, &total);
		const u32 nn = queue_advance(qs, req);
		cur = nxt;
		nxt = nn;
		I = In;
	}
}

__global__ __launch_bounds__(kPartBlock) void batch_part_slice_kernel(const BuildPlan* __restrict__ builds,
	const BatchItem* __restrict__ items, u32 first, u32 count, u32* __restrict__ counter)
{
	extern __shared__ u64 lds_words[];
	__shared__ u32 pre[kMaxChunks];
	__shared__ u32 base[kMaxChunks];
	__shared__ u32 total_rows;
	__shared__ u32 qs[2];
	u32 cur, nxt;
	queue_begin(counter, qs, cur, nxt);
	BatchItem I = items[first + (cur < count ? cur : 0)];
	while (cur < count) {
		const u32 req = queue_request(counter);
		const BatchItem In = items[first + (nxt < count ? nxt : cur)];
		part_slice(builds[I.build], (int)I.r0, lds_words, pre, base, &total_rows);
		const u32 nn = queue_advance(qs, req);
		cur = nxt;
		nxt = nn;
		I = In;
	}
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

hipError_t launch_batch_part(bool slices, const BuildPlan* d_builds, const BatchItem* d_items, u32 first, u32 count,
	u32* d_counter, int grid, size_t slice_bytes, hipStream_t s)
{
	if (!count)
		return hipSuccess;
	if (slices)
		hipLaunchKernelGGL(batch_part_slice_kernel, dim3(grid), dim3(kPartBlock), slice_bytes, s, d_builds, d_items,
			first, count, d_counter);
	else
		hipLaunchKernelGGL(batch_part_scatter_kernel, dim3(grid), dim3(kPartBlock), 0, s, d_builds, d_items, first,
			count, d_counter);
	return hipGetLastError();
}

size_t part_slice_lds_bytes(const BuildPlan& B)
{
	if (B.hashed)
		return ((size_t)1 << B.pshift) * sizeof(u64);
	return (((size_t)1 << B.pshift) * (size_t)B.dir_w + 3) / 4 * 4;
}

hipError_t launch_part_build(const BuildPlan& B, const BuildPlan* dB, hipStream_t s)
{
	if (B.nchunks > 0)
		hipLaunchKernelGGL(part_scatter_kernel, dim3((unsigned)B.nchunks), dim3(kPartBlock), 0, s, dB);
	const size_t slice = part_slice_lds_bytes(B);
	hipLaunchKernelGGL(part_build_kernel, dim3((unsigned)B.nparts), dim3(kPartBlock), slice, s, dB);
	return hipGetLastError();
}

hipError_t launch_build(const BuildPlan* B, int grid, hipStream_t s)
{
	hipLaunchKernelGGL(build_kernel, dim3(grid), dim3(kBlock), 0, s, B);
	return hipGetLastError();
}

size_t groupby_lds_bytes(const Plan& P)
{
	size_t b = 0;
	if (P.mode == GB_REG)
		b = (size_t)(P.domain * P.nsums * kBlock * 12 + P.domain * kBlock * 4 + 16);
	else if (P.mode == GB_LDS)
		b = (size_t)(P.domain * P.nsums * 12 + P.domain * 4 + 16);
	else if (P.mode == GB_LDSHASH)
		b = (size_t)P.nsums * kLdsHashSlots * 12 + kLdsHashSlots * 8 + 16;
	else if (P.mode == GB_PART)
		b = (((size_t)P.gparts * 8 + 16 + 15) & ~(size_t)15) + (size_t)P.gstage * (size_t)P.grec;
	return (b + 15) & ~(size_t)15;
}

size_t join_lds_bytes(const JoinT& J)
{
	const u64 keys = (u64)(J.kmax - J.kmin) + 1;
	u64 b = 0;
	if (J.lds == LDS_DIR)
		b = keys * (u64)J.dir_w;
	else if (J.lds == LDS_BITS)
		b = (keys + 31) / 32 * 4;
	return (size_t)((b + 15) & ~15ull);
}

size_t pipeline_lds_bytes(const Plan& P)
{
	size_t end = groupby_lds_bytes(P);
	for (int j = 0; j < P.njoins; j++)
		if (P.joins[j].lds)
			end = std::max(end, (size_t)P.joins[j].lds_off + join_lds_bytes(P.joins[j]));
	return end;
}

__global__ __launch_bounds__(256) void dir_bits_kernel(const void* __restrict__ dir, int w, u64 nkeys,
	u32* __restrict__ bits)
{
	const u64 nwords = (nkeys + 31) / 32;
	for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < nwords; i += (u64)gridDim.x * blockDim.x) {
		const u32 empty = dir_empty(w);
		u32 word = 0;
		for (u32 b = 0; b < 32; b++) {
			const u64 k = i * 32 + b;
			if (k < nkeys && dir_load(dir, w, (u32)k) != empty)
				word |= 1u << b;
		}
		bits[i] = word;
	}
}

hipError_t launch_dir_bits(const void* dir, int w, u64 nkeys, u32* bits, hipStream_t s)
{
	const u64 nwords = (nkeys + 31) / 32;
	const unsigned grid = (unsigned)std::max<u64>(1, std::min<u64>((nwords + 255) / 256, 1024));
	hipLaunchKernelGGL(dir_bits_kernel, dim3(grid), dim3(256), 0, s, dir, w, nkeys, bits);
	return hipGetLastError();
}

size_t part_slice_bytes(const Plan& P)
{
	return ((size_t)1 << P.gshift) * ((size_t)P.nsums * 12 + 4) + 16;
}

hipError_t launch_merge(const Plan& P, const Plan* dP, int grid, hipStream_t s)
{
	if (P.mode == GB_PART) {
		hipLaunchKernelGGL(part_agg_kernel, dim3((unsigned)P.gparts), dim3(kPartBlockAgg), part_slice_bytes(P), s, dP);
		return hipGetLastError();
	}
	if ((P.mode != GB_REG && P.mode != GB_LDS) || P.flush)
		return hipSuccess;
	const u64 elems = P.domain * (u64)P.nsums + P.domain;
	hipLaunchKernelGGL(merge_kernel, dim3((unsigned)elems), dim3(kBlock), 0, s, dP, (u32)grid);
	return hipGetLastError();
}

hipError_t launch_pipeline(const Plan& P, const Plan* dP, int grid, hipStream_t s)
{
	const size_t lds = pipeline_lds_bytes(P);
#define VH_LAUNCH(M) return launch_pipeline_mode<M>(dP, grid, lds, s)
	switch (P.mode) {
	case GB_REG: VH_LAUNCH(GB_REG);
	case GB_LDS: VH_LAUNCH(GB_LDS);
	case GB_GLOBAL: VH_LAUNCH(GB_GLOBAL);
	case GB_LDSHASH: VH_LAUNCH(GB_LDSHASH);
	case GB_PART: VH_LAUNCH(GB_PART);
	default: VH_LAUNCH(GB_HASH);
	}
#undef VH_LAUNCH
}

} // namespace vh
