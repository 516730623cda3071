// Statement-level gfx950 kernels: one launch per VOILA primitive over whole
// columns, with the reference's calling convention (sel, inum, res, cols...)
// (kernel_generator.py:42-43): `sel == nullptr` means the dense range
// [0, inum); otherwise sel holds inum ascending i32 positions.  Maps write
// res at the same position; selections write a compacted, ascending list.
//
// Table handles: a chained hash table in device memory (IHashTable analog,
// runtime_struct.hpp:453-503) with rows addressed by "bucket handles"
// row + 1 (0 = empty / end of chain), heads[capacity], next[rows], hash[rows]
// and one 64-bit (or 128-bit) word per row and column.
#include "vh_device.hpp"
#include "vh_prims.hpp"

namespace vh {

static constexpr int kTile = (int)kSelTileRows;   // positions per selection tile
static constexpr int kThreads = 256;

__device__ __forceinline__ i64 pos_of(const int32_t* sel, i64 k)
{
	return sel ? (i64)sel[k] : k;
}

__device__ __forceinline__ void store1(void* p, int t, i64 i, i64 v)
{
	switch (t) {
	case VH_U8: case VH_I8: ((uint8_t*)p)[i] = (uint8_t)v; break;
	case VH_U16: case VH_I16: ((uint16_t*)p)[i] = (uint16_t)v; break;
	case VH_U32: case VH_I32: ((uint32_t*)p)[i] = (uint32_t)v; break;
	default: ((i64*)p)[i] = v; break;
	}
}

// value reduced to type t (C conversion: modulo 2^bits, then sign)
__device__ __forceinline__ i64 wrap(int t, i64 v)
{
	switch (t) {
	case VH_U8: return (uint8_t)v;
	case VH_I8: return (int8_t)v;
	case VH_U16: return (uint16_t)v;
	case VH_I16: return (int16_t)v;
	case VH_U32: return (uint32_t)v;
	case VH_I32: return (int32_t)v;
	default: return v;
	}
}

// ---- selection compaction: vec_seltrue / vec_selfalse ---------------------
// Pass 1 counts set flags per tile, pass 2 scans the tile counts, pass 3
// writes each tile's positions at its offset with a wave ballot + mbcnt
// prefix, so the output is ascending like the reference's.

__global__ __launch_bounds__(kThreads) void sel_count_kernel(const int32_t* sel, i64 inum,
	const uint8_t* pred, int want, u32* tile_counts)
{
	__shared__ u32 wsum[kThreads / kWave];
	const i64 t0 = (i64)blockIdx.x * kTile;
	u32 c = 0;
	for (int k = threadIdx.x; k < kTile; k += kThreads) {
		const i64 i = t0 + k;
		if (i < inum)
			c += (pred[pos_of(sel, i)] != 0) == (want != 0);
	}
	for (int m = 32; m >= 1; m >>= 1)
		c += __shfl_xor(c, m, kWave);
	if (threadIdx.x % kWave == 0)
		wsum[threadIdx.x / kWave] = c;
	__syncthreads();
	if (threadIdx.x == 0) {
		u32 s = 0;
		for (int w = 0; w < kThreads / kWave; w++)
			s += wsum[w];
		tile_counts[blockIdx.x] = s;
	}
}

// Exclusive scan of n tile counts in one block; total into *total.
__global__ __launch_bounds__(1024) void scan_counts_kernel(u32* counts, i64 n, int32_t* total)
{
	__shared__ u64 part[1024];
	const i64 per = (n + 1023) / 1024;
	const i64 b = (i64)threadIdx.x * per;
	u64 s = 0;
	for (i64 i = b; i < b + per && i < n; i++)
		s += counts[i];
	part[threadIdx.x] = s;
	__syncthreads();
	for (int off = 1; off < 1024; off <<= 1) {
		u64 v = threadIdx.x >= (unsigned)off ? part[threadIdx.x - off] : 0;
		__syncthreads();
		part[threadIdx.x] += v;
		__syncthreads();
	}
	u64 run = threadIdx.x ? part[threadIdx.x - 1] : 0;
	for (i64 i = b; i < b + per && i < n; i++) {
		const u32 c = counts[i];
		counts[i] = (u32)run;
		run += c;
	}
	if (threadIdx.x == 1023)
		*total = (int32_t)part[1023];
}

__global__ __launch_bounds__(kThreads) void sel_write_kernel(const int32_t* sel, i64 inum,
	const uint8_t* pred, int want, const u32* tile_offsets, int32_t* res)
{
	__shared__ u32 wcount[kThreads / kWave];
	const i64 t0 = (i64)blockIdx.x * kTile;
	u32 base = tile_offsets[blockIdx.x];
	const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
	for (int k0 = 0; k0 < kTile; k0 += kThreads) {
		const i64 i = t0 + k0 + threadIdx.x;
		bool f = false;
		i64 p = 0;
		if (i < inum) {
			p = pos_of(sel, i);
			f = (pred[p] != 0) == (want != 0);
		}
		const u64 mask = __ballot(f);
		const u32 before = __builtin_amdgcn_mbcnt_hi((u32)(mask >> 32), __builtin_amdgcn_mbcnt_lo((u32)mask, 0u));
		if (lane == 0)
			wcount[wave] = (u32)__popcll(mask);
		__syncthreads();
		u32 woff = 0, tot = 0;
		for (int w = 0; w < kThreads / kWave; w++) {
			if (w < wave)
				woff += wcount[w];
			tot += wcount[w];
		}
		if (f)
			res[base + woff + before] = (int32_t)p;
		base += tot;
		__syncthreads();
	}
}

// ---- maps ----------------------------------------------------------------

__global__ __launch_bounds__(kThreads) void cmp_kernel(int cmp, int t, const int32_t* sel, i64 inum,
	uint8_t* res, const void* a, const void* b, i64 c)
{
	for (i64 k = (i64)blockIdx.x * kThreads + threadIdx.x; k < inum; k += (i64)gridDim.x * kThreads) {
		const i64 i = pos_of(sel, k);
		const i64 x = load1(a, t, i);
		const i64 y = b ? load1(b, t, i) : c;
		res[i] = compare(cmp, t, x, y) ? 1 : 0;
	}
}

// vec_add/sub/mul: both operands cast to the result type first
// (kernel_generator.py:439-445).
__global__ __launch_bounds__(kThreads) void arith_kernel(int op, int rt, const int32_t* sel, i64 inum,
	void* res, int ta, const void* a, int tb, const void* b, i64 cb)
{
	for (i64 k = (i64)blockIdx.x * kThreads + threadIdx.x; k < inum; k += (i64)gridDim.x * kThreads) {
		const i64 i = pos_of(sel, k);
		const u64 x = (u64)wrap(rt, load1(a, ta, i));
		const u64 y = (u64)wrap(rt, b ? load1(b, tb, i) : cb);
		const u64 z = op == VH_E_ADD ? x + y : op == VH_E_SUB ? x - y : x * y;
		store1(res, rt, i, wrap(rt, (i64)z));
	}
}

__global__ __launch_bounds__(kThreads) void hash_kernel(int t, const int32_t* sel, i64 inum, u64* res,
	const u64* h, const void* col)
{
	for (i64 k = (i64)blockIdx.x * kThreads + threadIdx.x; k < inum; k += (i64)gridDim.x * kThreads) {
		const i64 i = pos_of(sel, k);
		const u64 v = (u64)load1(col, t, i);
		res[i] = h ? voila_rehash(h[i], v) : voila_hash(v);
	}
}

// ---- chained table ops -----------------------------------------------------

__global__ __launch_bounds__(kThreads) void bucket_lookup_kernel(const u64* heads, u64 mask,
	const int32_t* sel, i64 inum, u64* res, const u64* hash)
{
	for (i64 k = (i64)blockIdx.x * kThreads + threadIdx.x; k < inum; k += (i64)gridDim.x * kThreads) {
		const i64 i = pos_of(sel, k);
		res[i] = heads[hash[i] & mask];
	}
}

__global__ __launch_bounds__(kThreads) void bucket_next_kernel(const u64* next, u64 rows,
	const int32_t* sel, i64 inum, u64* res, const u64* bucket)
{
	for (i64 k = (i64)blockIdx.x * kThreads + threadIdx.x; k < inum; k += (i64)gridDim.x * kThreads) {
		const i64 i = pos_of(sel, k);
		const u64 b = bucket[i];
		res[i] = (b && b <= rows) ? next[b - 1] : 0;
	}
}

__global__ __launch_bounds__(kThreads) void check_kernel(const i64* colv, u64 rows, const int32_t* sel,
	i64 inum, uint8_t* res, const u64* bucket, int kt, const void* key)
{
	for (i64 k = (i64)blockIdx.x * kThreads + threadIdx.x; k < inum; k += (i64)gridDim.x * kThreads) {
		const i64 i = pos_of(sel, k);
		const u64 b = bucket[i];
		res[i] = (b && b <= rows && colv[b - 1] == load1(key, kt, i)) ? 1 : 0;
	}
}

__global__ __launch_bounds__(kThreads) void gather_kernel(const i64* colv, u64 rows, const int32_t* sel,
	i64 inum, void* res, int rt, const u64* bucket)
{
	for (i64 k = (i64)blockIdx.x * kThreads + threadIdx.x; k < inum; k += (i64)gridDim.x * kThreads) {
		const i64 i = pos_of(sel, k);
		const u64 b = bucket[i];
		if (b && b <= rows)
			store1(res, rt, i, colv[b - 1]);
	}
}

__global__ __launch_bounds__(kThreads) void scatter_kernel(i64* colv, u64 rows, const int32_t* sel,
	i64 inum, const u64* bucket, int vt, const void* val)
{
	for (i64 k = (i64)blockIdx.x * kThreads + threadIdx.x; k < inum; k += (i64)gridDim.x * kThreads) {
		const i64 i = pos_of(sel, k);
		const u64 b = bucket[i];
		if (b && b <= rows)
			colv[b - 1] = load1(val, vt, i);
	}
}

// aggr_sum into a 128-bit column (lo, hi interleaved) / aggr_count.  The
// lanes of a wave that share a row (few groups: most of them) first add
// their values in registers, one row at a time for up to kAggrRounds rows,
// and one lane adds the exact 128-bit sum into the workgroup's LDS table
// (kAggrSlots rows, slot = row mod kAggrSlots); lanes left over (a wave
// over many groups) add their own values there.  A row whose slot holds
// another row adds into the column with atomics; at the end the workgroup
// adds each slot into the column.
static constexpr int kAggrRounds = 16;
static constexpr int kAggrSlots = 64;

struct AggrSlots {
	u64 key[kAggrSlots];   // row + 1, 0 = free
	u64 lo[kAggrSlots];
	u64 hi[kAggrSlots];    // two's complement high word
};

__device__ __forceinline__ void aggr_add(AggrSlots& S, u64* col2, u64 b, u64 lo, i64 hi)
{
	const int s = (int)(b & (kAggrSlots - 1));
	u64 k = S.key[s];
	if (k == 0) {
		k = atomicCAS(&S.key[s], 0ull, b);
		if (k == 0)
			k = b;
	}
	if (k != b) {
		atomic_add_i128(&col2[2 * (b - 1)], (i64*)&col2[2 * (b - 1) + 1], lo, hi);
		return;
	}
	const u64 old = atomicAdd(&S.lo[s], lo);
	const u64 h = (u64)hi + ((old + lo < old) ? 1ull : 0ull);
	if (h)
		atomicAdd(&S.hi[s], h);
}

__global__ __launch_bounds__(kThreads) void aggr_kernel(u64* col2, u64 rows, const int32_t* sel, i64 inum,
	const u64* bucket, int vt, const void* val)
{
	__shared__ AggrSlots S;
	for (int s = threadIdx.x; s < kAggrSlots; s += kThreads) {
		S.key[s] = 0;
		S.lo[s] = 0;
		S.hi[s] = 0;
	}
	__syncthreads();
	const bool sgn = val ? (vt == VH_I8 || vt == VH_I16 || vt == VH_I32 || vt == VH_I64) : false;
	const int lane = threadIdx.x & (kWave - 1);
	// a workgroup-uniform trip count: every lane of a wave takes part in the
	// wave-wide steps
	for (i64 k0 = (i64)blockIdx.x * kThreads; k0 < inum; k0 += (i64)gridDim.x * kThreads) {
		const i64 k = k0 + threadIdx.x;
		u64 b = 0;
		i64 v = 0;
		if (k < inum) {
			const i64 i = pos_of(sel, k);
			b = bucket[i];
			if (b && b <= rows)
				v = val ? load1(val, vt, i) : 1;
			else
				b = 0;
		}
		const u64 lo = (u64)v;
		const i64 hi = (sgn && v < 0) ? -1 : 0;
		bool pending = b != 0;
#pragma unroll 1
		for (int round = 0; round < kAggrRounds; round++) {
			const u64 m = __ballot(pending);
			if (!m)
				break;
			const int leader = __ffsll((long long)m) - 1;
			const u64 lb = ((u64)(u32)__shfl((int)(u32)(b >> 32), leader, kWave) << 32) |
				(u64)(u32)__shfl((int)(u32)b, leader, kWave);
			const bool mine = pending && b == lb;
			u64 slo = mine ? lo : 0ull;
			i64 shi = mine ? hi : 0;
			wave_sum_i128(slo, shi);
			if (lane == leader)
				aggr_add(S, col2, lb, slo, shi);
			pending = pending && !mine;
		}
		if (pending)
			aggr_add(S, col2, b, lo, hi);
	}
	__syncthreads();
	for (int s = threadIdx.x; s < kAggrSlots; s += kThreads) {
		const u64 k = S.key[s];
		if (k)
			atomic_add_i128(&col2[2 * (k - 1)], (i64*)&col2[2 * (k - 1) + 1], S.lo[s], (i64)S.hi[s]);
	}
}

// bucket_insert (kernel_generator.py:470-519) in parallel form.  As in the
// reference, a tuple whose bucket head already points at a row created by
// this call (row >= *first_new, the row count when the call began) is a
// conflict: res = 0 and it retries in the next round of the group-by loop.
// Otherwise the tuple locks the head with a CAS from the value it read, takes
// a row, chains it to the old head and publishes it; a tuple that loses the
// CAS is a conflict too.  Rows are only taken by winners, so the table stays
// dense.  A full table sets *overflow and restores the head.
static constexpr u64 kLocked = ~0ull;
static constexpr int kInsertRounds = 8;

__global__ __launch_bounds__(kThreads) void bucket_insert_kernel(u64* heads, u64 mask, u64* next,
	u64* hashcol, u64 max_rows, u64* counters, u32* overflow, const int32_t* sel, i64 inum, u64* res,
	const u64* hash)
{
	const u64 first_new = counters[1];
	const int lane = threadIdx.x & (kWave - 1);
	// a workgroup-uniform trip count: every lane of a wave takes part in the
	// wave-wide steps
	for (i64 k0 = (i64)blockIdx.x * kThreads; k0 < inum; k0 += (i64)gridDim.x * kThreads) {
		const i64 k = k0 + threadIdx.x;
		bool go = k < inum;
		i64 i = 0;
		u64 h = 0, idx = 0;
		if (go) {
			i = pos_of(sel, k);
			h = hash[i];
			idx = h & mask;
			res[i] = 0;
		}
		// lanes of a wave with the same bucket head (few groups: most of
		// them): only the first tries it; the others would find it locked or
		// holding this call's row, a conflict either way
		bool pending = go;
#pragma unroll 1
		for (int round = 0; round < kInsertRounds; round++) {
			const u64 m = __ballot(pending);
			if (!m)
				break;
			const int leader = __ffsll((long long)m) - 1;
			const u64 li = ((u64)(u32)__shfl((int)(u32)(idx >> 32), leader, kWave) << 32) |
				(u64)(u32)__shfl((int)(u32)idx, leader, kWave);
			const bool same = pending && idx == li;
			if (same && lane != leader)
				go = false;
			pending = pending && !same;
		}
		if (!go)
			continue;
		const u64 head = __hip_atomic_load(&heads[idx], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
		if (head == kLocked || (head && head - 1 >= first_new))
			continue;
		if (atomicCAS(&heads[idx], head, kLocked) != head)
			continue;
		const u64 row = atomicAdd(&counters[0], 1ull);
		if (row >= max_rows) {
			atomicOr(overflow, 1u);
			__hip_atomic_store(&heads[idx], head, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
			continue;
		}
		next[row] = head;
		hashcol[row] = h;
		__hip_atomic_store(&heads[idx], row + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
		res[i] = row + 1;
	}
}

// counters[1] = counters[0]: start of a bucket_insert call
__global__ void insert_snapshot_kernel(u64* counters)
{
	counters[1] = counters[0];
}

__global__ __launch_bounds__(kThreads) void write_kernel(i64* dst, const int32_t* sel, i64 inum, int vt,
	const void* val)
{
	for (i64 k = (i64)blockIdx.x * kThreads + threadIdx.x; k < inum; k += (i64)gridDim.x * kThreads)
		dst[k] = load1(val, vt, pos_of(sel, k));
}

__global__ __launch_bounds__(kThreads) void index_build_kernel(u64* heads, u64 mask, u64* next,
	const u64* hashcol, u64 rows)
{
	for (u64 r = (u64)blockIdx.x * kThreads + threadIdx.x; r < rows; r += (u64)gridDim.x * kThreads) {
		const u64 old = atomicExch(&heads[hashcol[r] & mask], r + 1);
		next[r] = old;
	}
}

// ---- launchers -----------------------------------------------------------

static unsigned grid_for(i64 n)
{
	i64 g = (n + kThreads - 1) / kThreads;
	if (g < 1) g = 1;
	if (g > 8192) g = 8192;
	return (unsigned)g;
}

i64 sel_tiles(i64 inum) { return (inum + kTile - 1) / kTile; }

hipError_t launch_select(const int32_t* sel, i64 inum, const uint8_t* pred, int want, int32_t* res,
	u32* tile_scratch, int32_t* d_count, hipStream_t s)
{
	const i64 tiles = sel_tiles(inum);
	if (tiles == 0) {
		hipLaunchKernelGGL(scan_counts_kernel, dim3(1), dim3(1024), 0, s, tile_scratch, (i64)0, d_count);
		return hipGetLastError();
	}
	hipLaunchKernelGGL(sel_count_kernel, dim3((unsigned)tiles), dim3(kThreads), 0, s, sel, inum, pred, want, tile_scratch);
	hipLaunchKernelGGL(scan_counts_kernel, dim3(1), dim3(1024), 0, s, tile_scratch, tiles, d_count);
	hipLaunchKernelGGL(sel_write_kernel, dim3((unsigned)tiles), dim3(kThreads), 0, s, sel, inum, pred, want, tile_scratch, res);
	return hipGetLastError();
}

hipError_t launch_cmp(int cmp, int t, const int32_t* sel, i64 inum, uint8_t* res, const void* a,
	const void* b, i64 c, hipStream_t s)
{
	if (inum <= 0) return hipSuccess;
	hipLaunchKernelGGL(cmp_kernel, dim3(grid_for(inum)), dim3(kThreads), 0, s, cmp, t, sel, inum, res, a, b, c);
	return hipGetLastError();
}

hipError_t launch_arith(int op, int rt, const int32_t* sel, i64 inum, void* res, int ta, const void* a,
	int tb, const void* b, i64 cb, hipStream_t s)
{
	if (inum <= 0) return hipSuccess;
	hipLaunchKernelGGL(arith_kernel, dim3(grid_for(inum)), dim3(kThreads), 0, s, op, rt, sel, inum, res, ta, a, tb, b, cb);
	return hipGetLastError();
}

hipError_t launch_hash(int t, const int32_t* sel, i64 inum, u64* res, const u64* h, const void* col, hipStream_t s)
{
	if (inum <= 0) return hipSuccess;
	hipLaunchKernelGGL(hash_kernel, dim3(grid_for(inum)), dim3(kThreads), 0, s, t, sel, inum, res, h, col);
	return hipGetLastError();
}

hipError_t launch_bucket_lookup(const u64* heads, u64 mask, const int32_t* sel, i64 inum, u64* res,
	const u64* hash, hipStream_t s)
{
	if (inum <= 0) return hipSuccess;
	hipLaunchKernelGGL(bucket_lookup_kernel, dim3(grid_for(inum)), dim3(kThreads), 0, s, heads, mask, sel, inum, res, hash);
	return hipGetLastError();
}

hipError_t launch_bucket_next(const u64* next, u64 rows, const int32_t* sel, i64 inum, u64* res,
	const u64* bucket, hipStream_t s)
{
	if (inum <= 0) return hipSuccess;
	hipLaunchKernelGGL(bucket_next_kernel, dim3(grid_for(inum)), dim3(kThreads), 0, s, next, rows, sel, inum, res, bucket);
	return hipGetLastError();
}

hipError_t launch_check(const i64* colv, u64 rows, const int32_t* sel, i64 inum, uint8_t* res,
	const u64* bucket, int kt, const void* key, hipStream_t s)
{
	if (inum <= 0) return hipSuccess;
	hipLaunchKernelGGL(check_kernel, dim3(grid_for(inum)), dim3(kThreads), 0, s, colv, rows, sel, inum, res, bucket, kt, key);
	return hipGetLastError();
}

hipError_t launch_gather(const i64* colv, u64 rows, const int32_t* sel, i64 inum, void* res, int rt,
	const u64* bucket, hipStream_t s)
{
	if (inum <= 0) return hipSuccess;
	hipLaunchKernelGGL(gather_kernel, dim3(grid_for(inum)), dim3(kThreads), 0, s, colv, rows, sel, inum, res, rt, bucket);
	return hipGetLastError();
}

hipError_t launch_scatter(i64* colv, u64 rows, const int32_t* sel, i64 inum, const u64* bucket, int vt,
	const void* val, hipStream_t s)
{
	if (inum <= 0) return hipSuccess;
	hipLaunchKernelGGL(scatter_kernel, dim3(grid_for(inum)), dim3(kThreads), 0, s, colv, rows, sel, inum, bucket, vt, val);
	return hipGetLastError();
}

hipError_t launch_aggr(u64* col2, u64 rows, const int32_t* sel, i64 inum, const u64* bucket, int vt,
	const void* val, hipStream_t s)
{
	if (inum <= 0) return hipSuccess;
	hipLaunchKernelGGL(aggr_kernel, dim3(grid_for(inum)), dim3(kThreads), 0, s, col2, rows, sel, inum, bucket, vt, val);
	return hipGetLastError();
}

hipError_t launch_bucket_insert(u64* heads, u64 mask, u64* next, u64* hashcol, u64 max_rows, u64* counters,
	u32* overflow, const int32_t* sel, i64 inum, u64* res, const u64* hash, hipStream_t s)
{
	if (inum <= 0) return hipSuccess;
	hipLaunchKernelGGL(insert_snapshot_kernel, dim3(1), dim3(1), 0, s, counters);
	hipLaunchKernelGGL(bucket_insert_kernel, dim3(grid_for(inum)), dim3(kThreads), 0, s, heads, mask, next,
		hashcol, max_rows, counters, overflow, sel, inum, res, hash);
	return hipGetLastError();
}

hipError_t launch_write(i64* dst, const int32_t* sel, i64 inum, int vt, const void* val, hipStream_t s)
{
	if (inum <= 0) return hipSuccess;
	hipLaunchKernelGGL(write_kernel, dim3(grid_for(inum)), dim3(kThreads), 0, s, dst, sel, inum, vt, val);
	return hipGetLastError();
}

hipError_t launch_index_build(u64* heads, u64 mask, u64* next, const u64* hashcol, u64 rows, hipStream_t s)
{
	if (!rows) return hipSuccess;
	hipLaunchKernelGGL(index_build_kernel, dim3(grid_for((i64)rows)), dim3(kThreads), 0, s, heads, mask, next, hashcol, rows);
	return hipGetLastError();
}

} // namespace vh
