// Device functions of the fused scan -> select -> join* -> group-by path,
// shared by the one-pipeline kernel (vh_pipeline.hip) and the persistent
// batched kernel (vh_batch.hip).
//
// Work is done chunk-at-a-time: a lane owns U rows of a tile (rows
// base + u * blockDim + lane, so every load instruction of a wave reads 64
// consecutive elements).  Each column is loaded for all U rows back to back,
// each join issues the first slot load of all U rows before resolving any of
// them, so a lane keeps U independent memory requests in flight.  This is
// VOILA's vector-at-a-time evaluation (a 1024-tuple vector per primitive
// call) moved into registers.
#pragma once

#include "vh_device.hpp"
#include "vh_plan.hpp"

namespace vh {

static constexpr u64 kEmpty = ~0ull;
static constexpr int kU = 4;              // rows per lane per chunk
// Longest linear-probe sequence a kernel follows.  Tables stay below 70% load
// (calc_num_buckets), where sequences are a few slots long; a longer one
// means a damaged table and is reported (error bit 32) instead of followed.
static constexpr u64 kMaxProbe = 1ull << 16;

// Next slot of a linear probe sequence that wraps within aligned blocks of
// smask + 1 slots (the whole table, or one partition's sub-table).
__device__ __forceinline__ u64 hnext(u64 h, u64 smask)
{
	return (h & ~smask) | ((h + 1) & smask);
}

// The join flags word of a launch, read once from the builds' flags.
__device__ __forceinline__ u32 join_flags(const Plan& P)
{
	u32 m = 0;
	for (int j = 0; j < P.njoins; j++) {
		if (*P.joins[j].dups)
			m |= 1u << j;
		if (P.joins[j].hashed && *P.joins[j].spill)
			m |= 1u << (kSpillBit + j);
	}
	return m;
}

// --------------------------------------------------------------------------
// direct join tables: dir[key - kmin] = payload - pmin in w = 1, 2 or 4 bytes
// --------------------------------------------------------------------------

__device__ __forceinline__ u32 dir_empty(int w)
{
	return w == 1 ? 0xFFu : (w == 2 ? 0xFFFFu : 0xFFFFFFFFu);
}

__device__ __forceinline__ u32 dir_load(const void* dir, int w, u32 off)
{
	if (w == 1)
		return gload((const uint8_t*)dir, off);
	if (w == 2)
		return gload((const uint16_t*)dir, off);
	return gload((const u32*)dir, off);
}

// Entry `off` of a dir[] slice held in LDS.
__device__ __forceinline__ u32 dir_load_lds(const u32* slice, int w, u32 off)
{
	if (w == 1)
		return ((const uint8_t*)slice)[off];
	if (w == 2)
		return ((const uint16_t*)slice)[off];
	return slice[off];
}

// The 32-bit word holding key offset `off`, and the bit mask of its entry.
__device__ __forceinline__ u32* dir_word(void* dir, int w, u32 off, u32& mask, u32& shift)
{
	const u64 byte = (u64)off * (u64)w;
	shift = (u32)(byte & 3) * 8;
	mask = (w == 4 ? 0xFFFFFFFFu : ((1u << (8 * w)) - 1u)) << shift;
	return (u32*)((char*)dir + (byte & ~3ull));
}

// Claiming an entry takes one atomic AND, no retry loop: entries start
// all-ones and a payload is never all-ones, so AND-ing (payload in the entry,
// ones elsewhere) into the word returns all-ones in the entry to exactly one
// row, the first.  A later row with the same key sees a value other than
// all-ones, knows it is a repeat, and ORs back the value it saw: its AND can
// only have cleared bits of the first row's payload, and every bit cleared
// while set is set again by the row whose AND cleared it, after that AND.
// After the launch every entry holds its first row's payload exactly.
// Conjunctive predicates of one row.
__device__ __forceinline__ bool preds_pass_row(const Pred* preds, int npreds, i64 row)
{
	bool ok = true;
#pragma unroll
	for (int p = 0; p < VH_MAX_PREDS; p++) {
		if (p < npreds) {
			const Pred& Q = preds[p];
			ok = ok && compare(Q.cmp, Q.r.type, load1(Q.r.ptr, Q.r.type, row), Q.value);
		}
	}
	return ok;
}

// Key offset and packed payload word of build row `row`.  ok: the row passed
// the build predicates; cleared, with error bit 4, when a key column lies
// outside its declared range.  A multi-column key packs each column's offset
// at its bit shift, as the probe does (probe_key_off).
__device__ __forceinline__ void build_row(const BuildPlan& B, i64 row, bool& ok, u32& off, u32& pay)
{
	if (B.nkeys <= 1) {
		const i64 key = ok ? load1(B.key, B.key_type, row) : B.kmin;
		if (ok && (key < B.kmin || key > B.kmax)) {
			gatomic_or(B.error, 4u);
			ok = false;
		}
		off = (u32)(key - B.kmin);
	} else {
		u32 o = 0;
#pragma unroll
		for (int c = 0; c < VH_MAX_JOIN_KEYS; c++) {
			if (c < B.nkeys) {
				const i64 v = ok ? load1(B.keyc[c], B.keyc_type[c], row) : B.kminc[c];
				const u64 d = (u64)(v - B.kminc[c]);
				if (ok && d > B.krangec[c]) {
					gatomic_or(B.error, 4u);
					ok = false;
				}
				o |= (u32)d << B.kshiftc[c];
			}
		}
		off = ok ? o : 0u;
	}
	if (B.npay <= 1) {
		pay = (ok && B.npay) ? (u32)(load1(B.pay, B.pay_type, row) - B.pmin) : 0u;
	} else {
		u32 p = 0;
#pragma unroll
		for (int k = 0; k < VH_MAX_PAYLOAD; k++)
			if (k < B.npay && ok)
				p |= (u32)(load1(B.payc[k], B.payc_type[k], row) - B.pminc[k]) << B.pshiftc[k];
		pay = p;
	}
}

// --------------------------------------------------------------------------
// chunk state
// --------------------------------------------------------------------------

struct Chunk {
	i64 row[kU];
	bool alive[kU];
	u32 pay[VH_MAX_JOINS][kU];   // payload - pmin of each join's match
};

// Typed loads of one column for the U rows of a chunk (exec-masked by
// `alive`, so rows that no longer qualify fetch nothing).
__device__ __forceinline__ void load_col_u(const void* p, int t, const Chunk& C, i64 out[kU])
{
#define VH_LOADU(T)                                                   \
	_Pragma("unroll") for (int u = 0; u < kU; u++)                    \
		out[u] = C.alive[u] ? (i64)gload((const T*)p, C.row[u]) : 0;
	switch (t) {
	case VH_U8: VH_LOADU(uint8_t); break;
	case VH_I8: VH_LOADU(int8_t); break;
	case VH_U16: VH_LOADU(uint16_t); break;
	case VH_I16: VH_LOADU(int16_t); break;
	case VH_U32: VH_LOADU(uint32_t); break;
	case VH_I32: VH_LOADU(int32_t); break;
	default: VH_LOADU(i64); break;
	}
#undef VH_LOADU
}

__device__ __forceinline__ void role_u(const Plan& P, const Role& r, const Chunk& C, i64 out[kU])
{
	switch (r.kind) {
	case ROLE_COL:
		load_col_u(r.ptr, r.type, C, out);
		break;
	case ROLE_PAY: {
		// payload `item` of the join, packed in the entry's payload word
		const JoinT& J = P.joins[r.slot];
		const i64 pmin = J.pminc[r.item];
		const int sh = J.pshiftc[r.item];
		const u32 m = J.pmaskc[r.item];
		switch (r.slot) {
#define VH_PAYU(J) _Pragma("unroll") for (int u = 0; u < kU; u++) out[u] = (i64)((C.pay[J][u] >> sh) & m) + pmin;
		case 0: VH_PAYU(0); break;
		case 1: VH_PAYU(1); break;
		case 2: VH_PAYU(2); break;
		default: VH_PAYU(3); break;
#undef VH_PAYU
		}
		break;
	}
	case ROLE_CONST:
#pragma unroll
		for (int u = 0; u < kU; u++)
			out[u] = r.c;
		break;
	default:
#pragma unroll
		for (int u = 0; u < kU; u++)
			out[u] = 0;
		break;
	}
}

// Key offsets of the chunk's rows for join J (key - kmin, or a multi-column
// key's packed offsets); rows whose key lies outside its range die.
__device__ __forceinline__ void probe_key_off(const Plan& P, const JoinT& J, Chunk& C, u32 off[kU])
{
	if (J.nkeys <= 1) {
		i64 key[kU];
		role_u(P, J.key, C, key);
#pragma unroll
		for (int u = 0; u < kU; u++) {
			if (C.alive[u] && (key[u] < J.kmin || key[u] > J.kmax))
				C.alive[u] = false;
			off[u] = (u32)(key[u] - J.kmin);
		}
		return;
	}
#pragma unroll
	for (int u = 0; u < kU; u++)
		off[u] = 0;
#pragma unroll 1
	for (int c = 0; c < VH_MAX_JOIN_KEYS; c++) {
		if (c < J.nkeys) {
			i64 v[kU];
			role_u(P, J.keyc[c], C, v);
#pragma unroll
			for (int u = 0; u < kU; u++) {
				const u64 d = (u64)(v[u] - J.kminc[c]);
				if (d > J.krangec[c])
					C.alive[u] = false;
				off[u] |= (u32)d << J.kshiftc[c];
			}
		}
	}
#pragma unroll
	for (int u = 0; u < kU; u++)
		off[u] = C.alive[u] ? off[u] : 0u;
}

__device__ __forceinline__ void eval_sum_u(const Plan& P, const SumArg& s, const Chunk& C, u64 prod[kU])
{
#pragma unroll 1
	for (int t = 0; t < 3; t++) {
		if (t < s.nterms) {
			const Term& T = s.t[t];
			i64 v[kU];
#pragma unroll
			for (int u = 0; u < kU; u++)
				v[u] = T.c;
			if (T.sa) {
				i64 a[kU];
				role_u(P, T.a, C, a);
#pragma unroll
				for (int u = 0; u < kU; u++)
					v[u] += T.sa > 0 ? a[u] : -a[u];
			}
			if (T.sb) {
				i64 b[kU];
				role_u(P, T.b, C, b);
#pragma unroll
				for (int u = 0; u < kU; u++)
					v[u] += T.sb > 0 ? b[u] : -b[u];
			}
#pragma unroll
			for (int u = 0; u < kU; u++)
				prod[u] = t == 0 ? (u64)v[u] : prod[u] * (u64)v[u];
		}
	}
}

// --------------------------------------------------------------------------
// aggregation state
// --------------------------------------------------------------------------

// LDS areas of the group-by modes (all per block):
//   GB_REG   lane-private accumulators: lo[(g*nsums+s)*bd + tid] (u64),
//            hi[...] (int, wide sums), cnt[g*bd + tid] (u32); no atomics
//   GB_LDS   shared dense table: lo[s*D+g], hi[s*D+g], cnt[g] with LDS atomics
//   GB_LDSHASH shared hash table of kLdsHashSlots over the dense index
struct Lds {
	u64* lo;
	int* hi;
	u32* cnt;
	u32* keys;
	// GB_PART: this region's tuple space (rows [rbase, rbase + rcap)) and
	// its index in the [region][partition] tables; cnt = per-partition
	// counts, keys = per-partition cursors, hi = the region's tuple count,
	// stage = its first Plan::gstage tuples
	i64 rbase;
	u32 rcap;
	u32 region;
	char* stage;
};

// One slot per active lane of the wave from a counter in memory, with one
// atomic per wave: lanes are ranked by mbcnt over the active mask.
__device__ __forceinline__ u32 wave_append(u32* ctr)
{
	const u64 mask = __ballot(1);
	const u32 rank = __builtin_amdgcn_mbcnt_hi((u32)(mask >> 32), __builtin_amdgcn_mbcnt_lo((u32)mask, 0u));
	const int leader = __ffsll((long long)mask) - 1;
	u32 base = 0;
	if ((int)(threadIdx.x & (kWave - 1)) == leader)
		base = atomicAdd(ctr, (u32)__popcll(mask));
	base = __shfl(base, leader, kWave);
	return base + rank;
}

// In-place exclusive prefix sum of a[0, n) in LDS (n <= kWave * 64) by the
// first wave: each lane sums a contiguous segment, a wave scan of the
// segment sums, then each lane rewrites its segment.  One barrier; every
// thread gets the total (through *total, in LDS).
// Rows laid out as runs (one per region or chunk): pre[r] = exclusive prefix
// of the run lengths over n >= 1 runs (pre[0] = 0, non-decreasing).  For each
// row ordinal i[v], the last run r[v] with pre[r[v]] <= i[v], by binary
// lifting over a fixed step sequence.  Every step reads one clamped, always
// valid word per row and updates the run with a select, so the V reads of a
// step are issued together and waited for once (a data-dependent search
// loop per row, or a read inside a branch, waits once per read).
template <int V>
__device__ __forceinline__ void runs_of(const u32* pre, int n, const u32 (&i)[V], int (&r)[V])
{
#pragma unroll
	for (int v = 0; v < V; v++)
		r[v] = 0;
	for (int step = n > 1 ? 1 << (31 - __builtin_clz((unsigned)(n - 1))) : 0; step > 0; step >>= 1) {
		u32 pm[V];
#pragma unroll
		for (int v = 0; v < V; v++)
			pm[v] = pre[min(r[v] + step, n - 1)];
#pragma unroll
		for (int v = 0; v < V; v++)
			r[v] += (r[v] + step < n && pm[v] <= i[v]) ? step : 0;
	}
}

__device__ __forceinline__ u32 block_exclusive_scan(u32* a, int n, u32* total)
{
	if (threadIdx.x < (unsigned)kWave) {
		const int lane = threadIdx.x;
		const int per = (n + kWave - 1) / kWave;
		const int b0 = lane * per;
		u32 t = 0;
		for (int i = b0; i < b0 + per && i < n; i++)
			t += a[i];
		u32 x = t;
		for (int d = 1; d < kWave; d <<= 1) {
			const u32 y = __shfl_up(x, d, kWave);
			if (lane >= d)
				x += y;
		}
		u32 run = x - t;
		for (int i = b0; i < b0 + per && i < n; i++) {
			const u32 v = a[i];
			a[i] = run;
			run += v;
		}
		if (lane == kWave - 1)
			*total = x;
	}
	__syncthreads();
	return *total;
}

// Where a row's group lives: a dense index g (GB_REG / GB_LDS / GB_GLOBAL),
// an LDS hash slot (GB_LDSHASH; -1 = LDS full, use the global arrays at g)
// or a global hash slot (GB_HASH).
template <int MODE>
__device__ __forceinline__ i64 locate(const Plan& P, const Lds& L, u64 g)
{
	if (MODE == GB_LDSHASH) {
		const u32 key = (u32)g;
		u32 slot = (u32)voila_hash((u64)key) & (kLdsHashSlots - 1);
		for (int n = 0; n < kLdsHashProbes; n++) {
			const u32 k = L.keys[slot];
			if (k == key)
				return slot;
			if (k == 0xFFFFFFFFu) {
				const u32 old = atomicCAS(&L.keys[slot], 0xFFFFFFFFu, key);
				if (old == 0xFFFFFFFFu || old == key)
					return slot;
			}
			slot = (slot + 1) & (kLdsHashSlots - 1);
		}
		return -1;
	}
	if (MODE == GB_HASH) {
		u64 h = voila_hash(g) & P.hmask;
		for (u64 n = 0; n <= P.hmask && n < kMaxProbe; n++) {
			const u64 k = gload(P.hkeys, (i64)h);
			if (k == g)
				return (i64)h;
			if (k == kEmpty) {
				const u64 old = gatomic_cas(&P.hkeys[h], kEmpty, g);
				if (old == kEmpty || old == g)
					return (i64)h;
			}
			h = (h + 1) & P.hmask;
		}
		gatomic_or(P.error, 1u);   // group hash table full
		return -1;
	}
	return (i64)g;
}

__device__ __forceinline__ void add_lds_exact(u64* lo, int* hi, u64 v, bool track, bool sgn)
{
	if (track) {
		const u64 o = atomicAdd(lo, v);
		const int h = (int)((o + v) < o) - (int)(sgn && (i64)v < 0);
		if (h)
			atomicAdd(hi, h);
	} else {
		atomicAdd(lo, v);
	}
}

__device__ __forceinline__ void add_global(u64* lo, i64* hi, u64 v, bool exact, bool sgn)
{
	if (exact)
		atomic_add_i128(lo, hi, v, (sgn && (i64)v < 0) ? -1 : 0);
	else
		gatomic_add(lo, v);
}

template <int MODE>
__device__ __forceinline__ void add_count(const Plan& P, const Lds& L, u64 g, i64 slot)
{
	if (MODE == GB_REG)
		L.cnt[g * blockDim.x + threadIdx.x] += 1u;
	else if (MODE == GB_LDS)
		atomicAdd(&L.cnt[g], 1u);
	else if (MODE == GB_LDSHASH && slot >= 0)
		atomicAdd(&L.cnt[slot], 1u);
	else if (MODE == GB_HASH)
		gatomic_add(&P.out_cnt[slot], 1ull);
	else
		gatomic_add(&P.out_cnt[g], 1ull);
}

template <int MODE>
__device__ __forceinline__ void add_sum(const Plan& P, const Lds& L, int s, u64 g, i64 slot, u64 v)
{
	const SumArg& S = P.sums[s];
	const bool track = P.wide && S.wide;
	const bool sgn = S.arg_signed;
	if (MODE == GB_REG) {
		const u64 idx = (g * P.nsums + s) * blockDim.x + threadIdx.x;
		const u64 o = L.lo[idx];
		const u64 n = o + v;
		L.lo[idx] = n;
		if (track)
			L.hi[idx] += (int)(n < o) - (int)(sgn && (i64)v < 0);
	} else if (MODE == GB_LDS) {
		const u64 idx = (u64)s * P.domain + g;
		add_lds_exact(&L.lo[idx], &L.hi[idx], v, track, sgn);
	} else if (MODE == GB_LDSHASH && slot >= 0) {
		const u64 idx = (u64)s * kLdsHashSlots + (u64)slot;
		add_lds_exact(&L.lo[idx], &L.hi[idx], v, track, sgn);
	} else if (MODE == GB_HASH) {
		const u64 idx = (u64)s * (P.hmask + 1) + (u64)slot;
		add_global(&P.out_lo[idx], &P.out_hi[idx], v, P.wide && S.res128, sgn);
	} else {
		const u64 idx = (u64)s * P.domain + g;
		add_global(&P.out_lo[idx], &P.out_hi[idx], v, P.wide && S.res128, sgn);
	}
}

// GB_PART tuple records (Plan::grec bytes): the group index (u32) at byte
// 0, sum s's argument at SumArg::toff, in 4 bytes (SumArg::narrow,
// sign-extended when the argument is signed and read back) or 8.  Tuple t
// of a region (its slot) is in the LDS stage while t < gstage, else in
// gtup0 at the region's row t; each store goes through the pointer of its
// own address space (a generic pointer to either would compile to FLAT
// stores, which the compiler can only wait for together with every load in
// flight).
#define VH_LDS __attribute__((address_space(3)))
static constexpr u32 kPartNoSlot = 0xFFFFFFFFu;

__device__ __forceinline__ void part_store(const Plan& P, const Lds& L, u32 t, int off, u64 v, bool wide)
{
	if (t < (u32)P.gstage) {
		VH_LDS char* r = (VH_LDS char*)L.stage + t * (u32)P.grec + (u32)off;
		if (wide)
			*(VH_LDS u64*)r = v;
		else
			*(VH_LDS u32*)r = (u32)v;
	} else {
		VH_GLOBAL char* r = (VH_GLOBAL char*)P.gtup0 + (u64)(L.rbase + t) * (u64)P.grec + off;
		if (wide)
			*(VH_GLOBAL u64*)r = v;
		else
			*(VH_GLOBAL u32*)r = (u32)v;
	}
}

// Sum s's argument out of a tuple record held as words in registers (the
// word is picked by comparisons, not by indexing the array).
template <int N>
__device__ __forceinline__ u64 part_field(const Plan& P, const u64 (&wd)[N], int s)
{
	const int o = P.sums[s].toff;
	u64 word = 0;
#pragma unroll
	for (int w = 0; w < N; w++)
		if (w == (o >> 3))
			word = wd[w];
	if (P.sums[s].narrow) {
		const u32 x = (u32)(word >> ((o & 7) * 8));
		return P.sums[s].arg_signed ? (u64)(i64)(int)x : (u64)x;
	}
	return word;
}

// GB_PART, one tuple per live row: a slot t in the region's tuple space
// (one atomic per wave) and the group index there, then each sum's
// argument; the partition's count in LDS.  Rows beyond the space (JoinN
// fan-out) add into the global arrays (t = kPartNoSlot).
__device__ __forceinline__ void part_tuple_slot(const Plan& P, const Lds& L, u64 g, bool live, u32& t)
{
	t = kPartNoSlot;
	if (!live)
		return;
	const u32 pos = wave_append((u32*)L.hi);
	if (pos < L.rcap) {
		t = pos;
		part_store(P, L, t, 0, g, false);
		atomicAdd(&L.cnt[g >> P.gshift], 1u);
	} else {
		gatomic_add(&P.out_cnt[g], 1ull);
	}
}

__device__ __forceinline__ void part_tuple_value(const Plan& P, const Lds& L, u32 t, int s, u64 g, u64 v)
{
	if (t != kPartNoSlot) {
		part_store(P, L, t, P.sums[s].toff, v, !P.sums[s].narrow);
	} else {
		const SumArg& S = P.sums[s];
		add_global(&P.out_lo[(u64)s * P.domain + g], &P.out_hi[(u64)s * P.domain + g], v, P.wide && S.res128,
			S.arg_signed);
	}
}

__device__ __forceinline__ void part_consume_u(const Plan& P, const Lds& L, const Chunk& C, const u64 g[kU],
	const bool ok[kU])
{
	u32 t[kU];
	bool live[kU];
#pragma unroll
	for (int u = 0; u < kU; u++) {
		live[u] = C.alive[u];
		if (live[u] && (!ok[u] || g[u] >= P.domain)) {
			gatomic_or(P.error, 8u);   // group key outside its column's [min, max]
			live[u] = false;
		}
		part_tuple_slot(P, L, g[u], live[u], t[u]);
	}
#pragma unroll 1
	for (int s = 0; s < VH_MAX_AGGRS; s++) {
		if (s < P.nsums) {
			u64 v[kU];
			eval_sum_u(P, P.sums[s], C, v);
#pragma unroll
			for (int u = 0; u < kU; u++)
				if (live[u])
					part_tuple_value(P, L, t[u], s, g[u], v[u]);
		}
	}
}

// End of a GB_PART region: order its tuples by partition (LDS stage, then
// gtup0 -> gtup1, the same rows' positions; records are moved as grec / 8
// words) and record each partition's run.
__device__ __forceinline__ void part_finish(const Plan& P, const Lds& L)
{
	__syncthreads();
	const u32 n = (u32)L.hi[0] < L.rcap ? (u32)L.hi[0] : L.rcap;
	const int NP = P.gparts;
	VH_GLOBAL u32* end = (VH_GLOBAL u32*)P.gend + (u64)L.region * NP;
	for (int p = threadIdx.x; p < NP; p += blockDim.x)
		L.keys[p] = L.cnt[p];
	__syncthreads();
	block_exclusive_scan(L.keys, NP, (u32*)&L.hi[1]);
	for (int p = threadIdx.x; p < NP; p += blockDim.x)
		end[p] = L.keys[p] + L.cnt[p];
	__syncthreads();
	const int W = P.grec / 8;
	constexpr int kMaxW = (4 + 8 * VH_MAX_AGGRS + 7) / 8;
	const u32 staged = n < (u32)P.gstage ? n : (u32)P.gstage;
	for (u32 i = threadIdx.x; i < staged; i += blockDim.x) {
		const VH_LDS u64* src = (const VH_LDS u64*)((VH_LDS char*)L.stage + i * (u32)P.grec);
		u64 x[kMaxW];
#pragma unroll
		for (int w = 0; w < kMaxW; w++)
			x[w] = w < W ? src[w] : 0ull;
		const u32 pos = atomicAdd(&L.keys[(u32)x[0] >> P.gshift], 1u);
		VH_GLOBAL u64* dst = (VH_GLOBAL u64*)P.gtup1 + (u64)(L.rbase + pos) * W;
#pragma unroll
		for (int w = 0; w < kMaxW; w++)
			if (w < W)
				dst[w] = x[w];
	}
	// every word of a thread's V tuples is loaded before any is stored, so
	// the copy costs one memory round trip per V tuples
	constexpr int V = 4;
	for (u32 i0 = staged; i0 < n; i0 += V * blockDim.x) {
		u64 x[V][kMaxW];
		bool has[V];
#pragma unroll
		for (int v = 0; v < V; v++) {
			const u32 i = i0 + v * blockDim.x + threadIdx.x;
			has[v] = i < n;
			const VH_GLOBAL u64* src = (const VH_GLOBAL u64*)P.gtup0 + (u64)(L.rbase + (has[v] ? i : 0u)) * W;
#pragma unroll
			for (int w = 0; w < kMaxW; w++)
				x[v][w] = (has[v] && w < W) ? src[w] : 0ull;
		}
#pragma unroll
		for (int v = 0; v < V; v++) {
			if (!has[v])
				continue;
			const u32 pos = atomicAdd(&L.keys[(u32)x[v][0] >> P.gshift], 1u);
			VH_GLOBAL u64* dst = (VH_GLOBAL u64*)P.gtup1 + (u64)(L.rbase + pos) * W;
#pragma unroll
			for (int w = 0; w < kMaxW; w++)
				if (w < W)
					dst[w] = x[v][w];
		}
	}
	__syncthreads();
}

// GB_PART, one workgroup per partition p: aggregate the partition's runs of
// every region in an LDS slice of 2^gshift groups, then add the slice into
// out_* (this workgroup is the only writer of these groups).  A thread
// loads V tuples of at most KW words before adding any.
template <int V, int KW>
__device__ __forceinline__ void part_agg_t(const Plan& P, int p, u64* lds, u32* pre, u32* base, u32* total_slot)
{
	const u64 S = 1ull << P.gshift;
	const u64 g0 = (u64)p * S;
	const u64 ng = S < P.domain - g0 ? S : P.domain - g0;
	const int ns = P.nsums;
	u64* lo = lds;
	int* hi = (int*)(lo + (u64)ns * S);
	u32* cnt = (u32*)(hi + (u64)ns * S);
	for (u64 i = threadIdx.x; i < (u64)ns * S; i += blockDim.x) {
		lo[i] = 0;
		hi[i] = 0;
	}
	for (u64 i = threadIdx.x; i < S; i += blockDim.x)
		cnt[i] = 0;
	const int NP = P.gparts;
	const int R = (int)((P.rows + P.gregion_rows - 1) / P.gregion_rows);
	// region r's run of partition p: [end[r][p - 1], end[r][p]) of the
	// region (two neighbouring words, one line read per region); a thread's
	// eight regions are loaded before any is stored
	constexpr int RV = 8;
	for (int r0 = threadIdx.x; r0 < R; r0 += RV * blockDim.x) {
		u32 lo_e[RV], hi_e[RV];
#pragma unroll
		for (int k = 0; k < RV; k++) {
			const int r = r0 + k * (int)blockDim.x;
			const VH_GLOBAL u32* e = (const VH_GLOBAL u32*)P.gend + (u64)(r < R ? r : 0) * NP;
			lo_e[k] = (r < R && p) ? e[p - 1] : 0u;
			hi_e[k] = r < R ? e[p] : 0u;
		}
#pragma unroll
		for (int k = 0; k < RV; k++) {
			const int r = r0 + k * (int)blockDim.x;
			if (r < R) {
				pre[r] = hi_e[k] - lo_e[k];
				base[r] = lo_e[k];
			}
		}
	}
	__syncthreads();
	const u32 total = block_exclusive_scan(pre, R, total_slot);
	// tuple i of this partition (0 <= i < total) sits at base[r] + i for the
	// last region r with pre[r] <= i
	for (int r = threadIdx.x; r < R; r += blockDim.x)
		base[r] = (u32)((i64)r * P.gregion_rows) + base[r] - pre[r];
	__syncthreads();
	// every word of a thread's V tuples is loaded before the LDS adds
	const int W = P.grec / 8;
	const u32 idle = threadIdx.x % (u32)ng;   // ng >= 1: this partition's groups
	for (u32 i0 = 0; i0 < total; i0 += V * blockDim.x) {
		bool has[V];
		u64 wd[V][KW];
		u32 iv[V], at[V];
		int rv[V];
#pragma unroll
		for (int v = 0; v < V; v++)
			iv[v] = i0 + v * blockDim.x + threadIdx.x;
		runs_of<V>(pre, R, iv, rv);
#pragma unroll
		for (int v = 0; v < V; v++)
			at[v] = base[rv[v]];
#pragma unroll
		for (int v = 0; v < V; v++) {
			const u32 i = iv[v];
			has[v] = i < total;
			const VH_GLOBAL u64* t = (const VH_GLOBAL u64*)P.gtup1 + (u64)(has[v] ? at[v] + i : 0u) * W;
#pragma unroll
			for (int w = 0; w < KW; w++)
				wd[v][w] = (has[v] && w < W) ? t[w] : 0ull;
		}
		// every lane adds (0 into a slot of its own for a tuple past the
		// total), so each sum's V atomics are issued together and a tracked
		// sum waits once for its V old values, not once per tuple
		u64 li[V];
#pragma unroll
		for (int v = 0; v < V; v++)
			li[v] = has[v] ? (u64)((u32)wd[v][0] - (u32)g0) : (u64)idle;
#pragma unroll
		for (int v = 0; v < V; v++)
			atomicAdd(&cnt[li[v]], has[v] ? 1u : 0u);
#pragma unroll
		for (int s2 = 0; s2 < VH_MAX_AGGRS; s2++) {
			if (s2 < ns) {
				const SumArg& A = P.sums[s2];
				u64 x[V];
#pragma unroll
				for (int v = 0; v < V; v++)
					x[v] = has[v] ? part_field(P, wd[v], s2) : 0ull;
				if (P.wide && A.wide) {
					u64 o[V];
#pragma unroll
					for (int v = 0; v < V; v++)
						o[v] = atomicAdd(&lo[(u64)s2 * S + li[v]], x[v]);
#pragma unroll
					for (int v = 0; v < V; v++) {
						const int h = (int)((o[v] + x[v]) < o[v]) - (int)(A.arg_signed && (i64)x[v] < 0);
						if (h)
							atomicAdd(&hi[(u64)s2 * S + li[v]], h);
					}
				} else {
#pragma unroll
					for (int v = 0; v < V; v++)
						atomicAdd(&lo[(u64)s2 * S + li[v]], x[v]);
				}
			}
		}
	}
	__syncthreads();
	// Without repeated keys in a JoinN's build no row fanned out into the
	// global arrays during the probe: every group's count is stored (zero
	// for the groups absent) and the sums of the groups present, with
	// nothing read first, so the arrays need no clearing between runs.
	bool fanout = false;
	for (int j = 0; j < P.njoins; j++)
		fanout = fanout || (P.joins[j].variant == VH_JOINN && *(const VH_GLOBAL u32*)P.joins[j].dups);
	if (!fanout) {
		for (u64 li = threadIdx.x; li < ng; li += blockDim.x) {
			const u32 c = cnt[li];
			const u64 g = g0 + li;
			P.out_cnt[g] = c;
			if (!c)
				continue;
#pragma unroll
			for (int s2 = 0; s2 < VH_MAX_AGGRS; s2++) {
				if (s2 < ns) {
					const SumArg& A = P.sums[s2];
					const u64 l = lo[(u64)s2 * S + li];
					const u64 idx = (u64)s2 * P.domain + g;
					P.out_lo[idx] = l;
					P.out_hi[idx] = (P.wide && A.wide) ? (i64)hi[(u64)s2 * S + li] : (A.arg_signed ? (i64)l >> 63 : 0);
				}
			}
		}
		return;
	}
	for (u64 li = threadIdx.x; li < ng; li += blockDim.x) {
		const u32 c = cnt[li];
		if (!c)
			continue;
		const u64 g = g0 + li;
		// this workgroup is the only writer of group g here; the global words
		// may hold fan-out rows added during the probe, so read them all first
		u64 olo[VH_MAX_AGGRS];
		i64 ohi[VH_MAX_AGGRS];
		const u64 oc = P.out_cnt[g];
#pragma unroll
		for (int s2 = 0; s2 < VH_MAX_AGGRS; s2++) {
			olo[s2] = s2 < ns ? P.out_lo[(u64)s2 * P.domain + g] : 0ull;
			ohi[s2] = s2 < ns ? P.out_hi[(u64)s2 * P.domain + g] : 0;
		}
		P.out_cnt[g] = oc + c;
#pragma unroll
		for (int s2 = 0; s2 < VH_MAX_AGGRS; s2++) {
			if (s2 < ns) {
				const SumArg& A = P.sums[s2];
				const u64 l = lo[(u64)s2 * S + li];
				const i64 h = (P.wide && A.wide) ? (i64)hi[(u64)s2 * S + li] : (A.arg_signed ? (i64)l >> 63 : 0);
				const u64 idx = (u64)s2 * P.domain + g;
				P.out_lo[idx] = olo[s2] + l;
				P.out_hi[idx] = ohi[s2] + h + ((olo[s2] + l) < olo[s2] ? 1 : 0);
			}
		}
	}
}

// part_agg_t with the tuple width of the plan: records of up to 16 bytes
// (two words) keep twice the tuples in flight in the same registers.
template <int V>
__device__ __forceinline__ void part_agg(const Plan& P, int p, u64* lds, u32* pre, u32* base, u32* total_slot)
{
	if (P.grec <= 16)
		part_agg_t<2 * V, 2>(P, p, lds, pre, base, total_slot);
	else
		part_agg_t<V, (4 + 8 * VH_MAX_AGGRS + 7) / 8>(P, p, lds, pre, base, total_slot);
}

// Group index (dense modes) or packed key (hash mode) of the chunk's live
// rows, their group slots, then one sum at a time: its argument for all U
// rows, added to each row's group.
template <int MODE>
__device__ __forceinline__ void consume_u(const Plan& P, const Lds& L, const Chunk& C)
{
	u64 g[kU];
	bool ok[kU];
#pragma unroll
	for (int u = 0; u < kU; u++) {
		g[u] = 0;
		ok[u] = C.alive[u];
	}
#pragma unroll 1
	for (int k = 0; k < VH_MAX_KEYS; k++) {
		if (k < P.nkeys) {
			i64 kv[kU];
			role_u(P, P.keys[k].r, C, kv);
#pragma unroll
			for (int u = 0; u < kU; u++) {
				const u64 off = (u64)(kv[u] - P.keys[k].min);
				ok[u] = ok[u] && off <= P.keys[k].range;
				g[u] += MODE == GB_HASH ? (off << P.keys[k].shift) : off * P.keys[k].mult;
			}
		}
	}
	if (MODE == GB_PART) {
		part_consume_u(P, L, C, g, ok);
		return;
	}
	i64 slot[kU];
	bool live[kU];
#pragma unroll
	for (int u = 0; u < kU; u++) {
		live[u] = C.alive[u];
		slot[u] = -1;
		if (!live[u])
			continue;
		if (!ok[u] || (MODE != GB_HASH && g[u] >= P.domain)) {
			gatomic_or(P.error, 8u);   // group key outside its column's [min, max]
			live[u] = false;
			continue;
		}
		slot[u] = locate<MODE>(P, L, g[u]);
		if (MODE == GB_HASH && slot[u] < 0) {
			live[u] = false;
			continue;
		}
		add_count<MODE>(P, L, g[u], slot[u]);
	}
#pragma unroll 1
	for (int s = 0; s < VH_MAX_AGGRS; s++) {
		if (s < P.nsums) {
			u64 v[kU];
			eval_sum_u(P, P.sums[s], C, v);
#pragma unroll
			for (int u = 0; u < kU; u++)
				if (live[u])
					add_sum<MODE>(P, L, s, g[u], slot[u], v[u]);
		}
	}
}

// --------------------------------------------------------------------------
// one chunk: select -> joins -> group-by
// --------------------------------------------------------------------------

// Join J0 and everything after it for the chunk's live rows.  The first
// slot of every live row is in flight before any is resolved.  A JoinN whose
// build saw a repeated key fans out: the rows' further matches are taken in
// lockstep (the k-th match of every row in the k-th round), each round
// running the rest of the pipeline, until no row has another.  Further
// matches follow the first in its probe sequence (hash layout; then the
// overflow table after a spill) or sit in the overflow slots from
// voila_hash(key) on (direct layout).
template <int MODE, int J0>
__device__ __forceinline__ void joins_from(const Plan& P, const Lds& L, Chunk& C, u32 dups_mask)
{
	if constexpr (J0 >= VH_MAX_JOINS) {
		consume_u<MODE>(P, L, C);
	} else {
		if (J0 >= P.njoins) {
			consume_u<MODE>(P, L, C);
			return;
		}
		const JoinT& J = P.joins[J0];
		u64 h[kU];
		u32 off[kU];
		u32 ovf_first = 0;   // rows whose match came from the overflow table
		probe_key_off(P, J, C, off);
#pragma unroll
		for (int u = 0; u < kU; u++)
			h[u] = voila_hash((u64)off[u]) & J.mask;
		if (J.dir_w) {
			// direct layout: one load per row
			u32 v[kU];
			const u32 empty = dir_empty(J.dir_w);
#pragma unroll
			for (int u = 0; u < kU; u++)
				v[u] = C.alive[u] ? dir_load(J.dir, J.dir_w, off[u]) : empty;
#pragma unroll
			for (int u = 0; u < kU; u++) {
				if (v[u] == empty)
					C.alive[u] = false;
				else
					C.pay[J0][u] = v[u];
				h[u] = (h[u] - 1) & J.mask;   // the overflow walk starts at h + 1
			}
		} else {
			u64 s[kU];
#pragma unroll
			for (int u = 0; u < kU; u++)
				s[u] = C.alive[u] ? gload(J.slots, (i64)h[u]) : kEmpty;
#pragma unroll
			for (int u = 0; u < kU; u++) {
				if (!C.alive[u])
					continue;
				// a sub-table may be full (its other rows spilled): one full
				// cycle of it, smask + 1 slots, ends the walk
				const u64 lim = J.smask < kMaxProbe ? J.smask + 1 : kMaxProbe;
				u64 n = 0;
				while (s[u] != kEmpty && (u32)s[u] != off[u] && n < lim) {
					n++;
					h[u] = hnext(h[u], J.smask);
					s[u] = gload(J.slots, (i64)h[u]);
				}
				if (n >= kMaxProbe)
					gatomic_or(P.error, 32u);
				if ((s[u] == kEmpty || (u32)s[u] != off[u]) && ((dups_mask >> (kSpillBit + J0)) & 1u)) {
					// not in its sub-table: rows of a full sub-table are in
					// the overflow table
					ovf_first |= 1u << u;
					h[u] = voila_hash((u64)off[u]) & J.mask;
					s[u] = gload(J.ovf, (i64)h[u]);
					for (n = 0; s[u] != kEmpty && (u32)s[u] != off[u] && n < kMaxProbe; n++) {
						h[u] = (h[u] + 1) & J.mask;
						s[u] = gload(J.ovf, (i64)h[u]);
					}
					if (n >= kMaxProbe)
						gatomic_or(P.error, 32u);
				}
				if (s[u] == kEmpty || (u32)s[u] != off[u])
					C.alive[u] = false;
				else
					C.pay[J0][u] = (u32)(s[u] >> 32);
			}
		}
		if (!(J.variant == VH_JOINN && ((dups_mask >> J0) & 1u))) {
			joins_from<MODE, J0 + 1>(P, L, C, dups_mask);
			return;
		}
		// fan-out: the rows' state after this join, and where each one's
		// sequence continues
		const Chunk keep = C;
		bool walking[kU], ov[kU];
#pragma unroll
		for (int u = 0; u < kU; u++) {
			walking[u] = keep.alive[u];
			ov[u] = (ovf_first >> u) & 1u;
			h[u] = ov[u] ? (h[u] + 1) & J.mask : hnext(h[u], J.smask);
		}
		for (int round = 0;; round++) {
			joins_from<MODE, J0 + 1>(P, L, C, dups_mask);
			C = keep;
			bool any = false;
#pragma unroll
			for (int u = 0; u < kU; u++) {
				C.alive[u] = false;
				if (!walking[u])
					continue;
				const u64 h0 = voila_hash((u64)off[u]) & J.mask;   // where the sequence began
				for (u64 n = 0;; n++) {
					if (n >= kMaxProbe) {
						gatomic_or(P.error, 32u);
						walking[u] = false;
						break;
					}
					// a full sub-table has no empty slot: its sequence ends
					// where it began
					const u64 sv = (!ov[u] && J.hashed && h[u] == h0) ? kEmpty
						: gload(ov[u] ? J.ovf : J.slots, (i64)h[u]);
					if (sv == kEmpty) {
						if (!ov[u] && ((dups_mask >> (kSpillBit + J0)) & 1u)) {
							ov[u] = true;
							h[u] = h0;
							continue;
						}
						walking[u] = false;
						break;
					}
					h[u] = ov[u] ? (h[u] + 1) & J.mask : hnext(h[u], J.smask);
					if ((u32)sv == off[u]) {
						C.alive[u] = true;
						C.pay[J0][u] = (u32)(sv >> 32);
						any = true;
						break;
					}
				}
			}
			if (!any)
				break;
		}
	}
}

template <int MODE>
__device__ __forceinline__ void process_chunk(const Plan& P, const Lds& L, Chunk& C, u32 dups_mask)
{
	// select: conjunctive predicates, one column at a time
#pragma unroll 1
	for (int p = 0; p < VH_MAX_PREDS; p++) {
		if (p < P.npreds) {
			const Pred& Q = P.preds[p];
			i64 x[kU];
			load_col_u(Q.r.ptr, Q.r.type, C, x);
#pragma unroll
			for (int u = 0; u < kU; u++)
				C.alive[u] = C.alive[u] && compare(Q.cmp, Q.r.type, x[u], Q.value);
		}
	}
	joins_from<MODE, 0>(P, L, C, dups_mask);
}

// Rows [r0, r1) in chunks of kU * blockDim.
template <int MODE>
__device__ __forceinline__ void process_range(const Plan& P, const Lds& L, i64 r0, i64 r1, i64 tile_stride,
	u32 dups_mask)
{
	const i64 tile = (i64)kU * blockDim.x;
	for (i64 base = r0; base < r1; base += tile_stride) {
		Chunk C;
#pragma unroll
		for (int u = 0; u < kU; u++) {
			C.row[u] = base + (i64)u * blockDim.x + threadIdx.x;
			C.alive[u] = C.row[u] < r1 && C.row[u] < base + tile;
		}
		process_chunk<MODE>(P, L, C, dups_mask);
	}
}

// LDS layout of each mode (sizes in pipeline_lds_bytes).
template <int MODE>
__device__ __forceinline__ Lds lds_layout(const Plan& P, u64* lds)
{
	Lds L = {nullptr, nullptr, nullptr, nullptr};
	if (MODE == GB_REG) {
		const u64 n = P.domain * P.nsums * blockDim.x;
		L.lo = lds;
		L.hi = (int*)(lds + n);
		L.cnt = (u32*)(L.hi + n);
	} else if (MODE == GB_LDS) {
		const u64 n = P.domain * P.nsums;
		L.lo = lds;
		L.hi = (int*)(lds + n);
		L.cnt = (u32*)(L.hi + n);
	} else if (MODE == GB_LDSHASH) {
		const u64 n = (u64)P.nsums * kLdsHashSlots;
		L.lo = lds;
		L.hi = (int*)(lds + n);
		L.cnt = (u32*)(L.hi + n);
		L.keys = L.cnt + kLdsHashSlots;
	} else if (MODE == GB_PART) {
		L.cnt = (u32*)lds;
		L.keys = L.cnt + P.gparts;
		L.hi = (int*)(L.keys + P.gparts);   // [0] tuple count, [1] scan total
		L.stage = (char*)lds + (((u64)P.gparts * 8 + 16 + 15) & ~15ull);
	}
	return L;
}

template <int MODE>
__device__ __forceinline__ void lds_init(const Plan& P, const Lds& L)
{
	u64 n = 0, groups = 0;
	if (MODE == GB_REG) { n = P.domain * P.nsums * blockDim.x; groups = P.domain * blockDim.x; }
	else if (MODE == GB_LDS) { n = P.domain * P.nsums; groups = P.domain; }
	else if (MODE == GB_LDSHASH) { n = (u64)P.nsums * kLdsHashSlots; groups = kLdsHashSlots; }
	for (u64 i = threadIdx.x; i < n; i += blockDim.x) {
		L.lo[i] = 0;
		L.hi[i] = 0;
	}
	for (u64 i = threadIdx.x; i < groups; i += blockDim.x) {
		L.cnt[i] = 0;
		if (MODE == GB_LDSHASH)
			L.keys[i] = 0xFFFFFFFFu;
	}
	if (MODE == GB_PART) {
		for (int i = threadIdx.x; i < P.gparts; i += blockDim.x)
			L.cnt[i] = 0;
		if (threadIdx.x == 0)
			L.hi[0] = 0;
	}
}

// GB_LDSHASH: add this block's LDS groups to the global arrays.
__device__ __forceinline__ void ldshash_flush(const Plan& P, const Lds& L)
{
	for (int i = threadIdx.x; i < kLdsHashSlots; i += blockDim.x) {
		const u32 key = L.keys[i];
		if (key == 0xFFFFFFFFu)
			continue;
		gatomic_add(&P.out_cnt[key], (u64)L.cnt[i]);
		for (int s = 0; s < P.nsums; s++) {
			const SumArg& S = P.sums[s];
			const u64 lo = L.lo[s * kLdsHashSlots + i];
			const i64 hi = (P.wide && S.wide) ? (i64)L.hi[s * kLdsHashSlots + i]
				: (S.arg_signed ? (i64)lo >> 63 : 0);
			const u64 idx = (u64)s * P.domain + key;
			if (S.res128)
				atomic_add_i128(&P.out_lo[idx], &P.out_hi[idx], lo, hi);
			else
				gatomic_add(&P.out_lo[idx], lo);
		}
	}
}

// End of a one-pipeline launch: GB_LDSHASH adds its LDS groups to the
// global arrays; GB_REG / GB_LDS write the block's exact partials
// (element-major, see merge_kernel) for the merge launch that follows, or
// with P.flush add them into the global arrays with 128-bit atomics.
template <int MODE>
__device__ __forceinline__ void block_epilogue(const Plan& P, const Lds& L)
{
	if (MODE == GB_PART) {
		part_finish(P, L);
		return;
	}
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
	const u64 nb = gridDim.x;
	u64* mine = P.partials + blockIdx.x;     // element k of this block: mine[k * nb]
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
					if (P.flush) {
						if (lo || hi)
							atomic_add_i128(&P.out_lo[e], &P.out_hi[e], lo, hi);
					} else {
						mine[(2 * e) * nb] = lo;
						mine[(2 * e + 1) * nb] = (u64)hi;
					}
				}
			} else {
				const u64 g = (u64)e - D * P.nsums;
				u64 c = 0;
				for (u32 t = lane; t < bd; t += kWave)
					c += L.cnt[g * bd + t];
				c = wave_sum_u64(c);
				if (lane == 0) {
					if (!P.flush)
						mine[(2 * D * P.nsums + g) * nb] = c;
					else if (c)
						gatomic_add(&P.out_cnt[g], c);
				}
			}
		}
	} else {
		const u64 n = D * P.nsums;
		for (u64 i = threadIdx.x; i < n; i += blockDim.x) {
			const SumArg& S = P.sums[i / D];
			const u64 lo = L.lo[i];
			const i64 hi = (P.wide && S.wide) ? (i64)L.hi[i] : (S.arg_signed ? (i64)lo >> 63 : 0);
			if (P.flush) {
				if (lo || hi)
					atomic_add_i128(&P.out_lo[i], &P.out_hi[i], lo, hi);
			} else {
				mine[(2 * i) * nb] = lo;
				mine[(2 * i + 1) * nb] = (u64)hi;
			}
		}
		for (u64 i = threadIdx.x; i < D; i += blockDim.x) {
			if (!P.flush)
				mine[(2 * n + i) * nb] = L.cnt[i];
			else if (L.cnt[i])
				gatomic_add(&P.out_cnt[i], (u64)L.cnt[i]);
		}
	}
}

// GB_LDS in the batched kernel: add a work item's dense LDS groups to the
// global arrays.
__device__ __forceinline__ void lds_dense_flush(const Plan& P, const Lds& L)
{
	const u64 D = P.domain;
	for (u64 g = threadIdx.x; g < D; g += blockDim.x) {
		const u32 c = L.cnt[g];
		if (!c)
			continue;
		gatomic_add(&P.out_cnt[g], (u64)c);
		for (int s = 0; s < P.nsums; s++) {
			const SumArg& S = P.sums[s];
			const u64 lo = L.lo[s * D + g];
			const i64 hi = (P.wide && S.wide) ? (i64)L.hi[s * D + g] : (S.arg_signed ? (i64)lo >> 63 : 0);
			if (S.res128)
				atomic_add_i128(&P.out_lo[s * D + g], &P.out_hi[s * D + g], lo, hi);
			else
				gatomic_add(&P.out_lo[s * D + g], lo);
		}
	}
}

// --------------------------------------------------------------------------
// persistent work queues (batched launches)
// --------------------------------------------------------------------------

// A persistent workgroup of a build launch holds two items of its launch's
// queue at a time: while it works on the current one, the slot of the one
// after next is being requested (thread 0's atomic on the launch's counter)
// and the next one's descriptor is being loaded, so neither round trip
// stands between two items (batch builds of 66 M rows in 4,096-row chunks:
// scatter 513 -> 486 us, slices 304 -> 291 us; the probe and aggregation
// launches, whose items are large and uneven, keep one item at a time: two
// unbalanced their tails, 2.23 -> 2.35 ms and 0.93 -> 1.12 ms).  qs: two
// words of LDS.  A slot at or past the item count ends the workgroup's loop:
//   u32 cur, nxt;
//   queue_begin(counter, qs, cur, nxt);
//   BatchItem I = items[first + (cur < count ? cur : 0)];
//   while (cur < count) {
//       const u32 req = queue_request(counter);
//       const BatchItem In = items[first + (nxt < count ? nxt : cur)];
//       ... item I ...
//       const u32 nn = queue_advance(qs, req);
//       cur = nxt; nxt = nn; I = In;
//   }
__device__ __forceinline__ void queue_begin(u32* counter, u32* qs, u32& cur, u32& nxt)
{
	if (threadIdx.x == 0) {
		qs[0] = gatomic_add(counter, 1u);
		qs[1] = gatomic_add(counter, 1u);
	}
	__syncthreads();
	cur = qs[0];
	nxt = qs[1];
	__syncthreads();
}

__device__ __forceinline__ u32 queue_request(u32* counter)
{
	return threadIdx.x == 0 ? gatomic_add(counter, 1u) : 0u;
}

// After an item: thread 0's request, made at the item's start, becomes every
// thread's slot after next.
__device__ __forceinline__ u32 queue_advance(u32* qs, u32 req)
{
	__syncthreads();
	if (threadIdx.x == 0)
		qs[0] = req;
	__syncthreads();
	return qs[0];
}

// --------------------------------------------------------------------------
// radix-partitioned probe (Plan::radix)
// --------------------------------------------------------------------------

// Probe partition of key offset `off` in join J: the aligned block of
// 2^rpshift slots its hash falls in.
__device__ __forceinline__ u32 radix_part(const Plan& P, const JoinT& J, u32 off)
{
	return (u32)((voila_hash((u64)off) & J.mask) >> P.rpshift);
}

// First half of the flush of chunk c: the tuples staged in LDS have been
// counted per partition in rc (rparts words); the partitions' running ends
// go to rend's row c and their starts to rcur (rparts words), the cursors
// the tuples are then placed with.  rtot: one word of LDS.
__device__ __forceinline__ void radix_flush_begin(const Plan& P, u32 c, const u32* rc, u32* rcur, u32* rtot)
{
	__syncthreads();
	const int NP = P.rparts;
	for (int q = threadIdx.x; q < NP; q += blockDim.x)
		rcur[q] = rc[q];
	__syncthreads();
	block_exclusive_scan(rcur, NP, rtot);
	VH_GLOBAL u32* end = (VH_GLOBAL u32*)P.rend + (u64)c * (u64)NP;
	for (int q = threadIdx.x; q < NP; q += blockDim.x)
		end[q] = rcur[q] + rc[q];
	__syncthreads();
}

// Second half: the counts and the staged tuple count start again at zero.
__device__ __forceinline__ void radix_flush_end(const Plan& P, u32* rc, u32* rn)
{
	__syncthreads();
	for (int q = threadIdx.x; q < P.rparts; q += blockDim.x)
		rc[q] = 0;
	if (threadIdx.x == 0)
		*rn = 0;
	__syncthreads();
}

// Empty runs for the chunk ids [from, rchunks) no iteration of the launch
// used (one workgroup writes them).
__device__ __forceinline__ void radix_pad(const Plan& P, u32 from)
{
	VH_GLOBAL u32* end = (VH_GLOBAL u32*)P.rend;
	for (u64 i = (u64)from * (u64)P.rparts + threadIdx.x; i < (u64)P.rchunks * (u64)P.rparts; i += blockDim.x)
		end[i] = 0;
}

// Join launch, partition p: pre / base (rchunks words of LDS each) locate
// its run of every chunk (tuple i of the partition is at base[c] + i for the
// last chunk c with pre[c] <= i); returns the partition's tuple count.
__device__ __forceinline__ u32 radix_runs(const Plan& P, u32 p, u32* pre, u32* base, u32* total)
{
	const int RC = P.rchunks;
	const u64 NP = (u64)P.rparts;
	for (int c = threadIdx.x; c < RC; c += blockDim.x) {
		const VH_GLOBAL u32* e = (const VH_GLOBAL u32*)P.rend + (u64)c * NP;
		const u32 b = p ? e[p - 1] : 0u;
		pre[c] = e[p] - b;
		base[c] = b;
	}
	__syncthreads();
	const u32 t = block_exclusive_scan(pre, RC, total);
	for (int c = threadIdx.x; c < RC; c += blockDim.x)
		base[c] = (u32)((u64)c * (u64)P.rchunk) + base[c] - pre[c];
	__syncthreads();
	return t;
}

__device__ __forceinline__ u64 radix_tuple_index(const u32* pre, const u32* base, int R, u32 i)
{
	int lo = 0, hi = R - 1;
	while (lo < hi) {
		const int mid = (lo + hi + 1) >> 1;
		if (pre[mid] <= i)
			lo = mid;
		else
			hi = mid - 1;
	}
	return (u64)base[lo] + i;
}

// --------------------------------------------------------------------------
// build side
// --------------------------------------------------------------------------

// Build side of one chunk (U rows per lane): select, then insert (key - kmin,
// payload - pmin).  Replaces write_pos/write + bucket_build
// (relalg_translator.cpp:828-928, runtime_struct.cpp:1283-1324).  The first
// CAS of every row is issued before any collision is resolved, so U atomics
// are in flight.
__device__ __forceinline__ void build_chunk(const BuildPlan& B, i64 base, i64 end)
{
	i64 row[kU];
	bool ok[kU];
#pragma unroll
	for (int u = 0; u < kU; u++) {
		row[u] = base + (i64)u * blockDim.x + threadIdx.x;
		ok[u] = row[u] < end;
	}
#pragma unroll
	for (int p = 0; p < VH_MAX_PREDS; p++) {
		if (p < B.npreds) {
			const Pred& Q = B.preds[p];
#pragma unroll
			for (int u = 0; u < kU; u++)
				ok[u] = ok[u] && compare(Q.cmp, Q.r.type, load1(Q.r.ptr, Q.r.type, row[u]), Q.value);
		}
	}
	u64 val[kU], h[kU], old[kU];
	u32 off[kU];
#pragma unroll
	for (int u = 0; u < kU; u++) {
		u32 pay;
		build_row(B, row[u], ok[u], off[u], pay);
		val[u] = (u64)off[u] | ((u64)pay << 32);
		h[u] = voila_hash((u64)off[u]) & B.mask;
	}
	if (B.dir_w) {
		// direct layout: claim dir[off]; a key seen before goes on to the
		// overflow slots below
		u32* word[kU];
		u32 mask[kU], bits[kU], seen[kU];
#pragma unroll
		for (int u = 0; u < kU; u++) {
			u32 shift;
			word[u] = dir_word(B.dir, B.dir_w, off[u], mask[u], shift);
			bits[u] = ((u32)(val[u] >> 32) << shift) & mask[u];
		}
#pragma unroll
		for (int u = 0; u < kU; u++)
			seen[u] = ok[u] ? gatomic_and(word[u], bits[u] | ~mask[u]) : 0xFFFFFFFFu;
		bool spill = false;
#pragma unroll
		for (int u = 0; u < kU; u++) {
			if (ok[u] && (seen[u] & mask[u]) != mask[u])
				gatomic_or(word[u], seen[u] & mask[u]);   // a repeated key: restore
			else
				ok[u] = false;                          // first row of its key
			spill = spill || ok[u];
		}
		if (!spill)
			return;
		*B.dups = 1u;
	}
#pragma unroll
	for (int u = 0; u < kU; u++)
		old[u] = ok[u] ? gatomic_cas(&B.slots[h[u]], kEmpty, val[u]) : kEmpty;
#pragma unroll
	for (int u = 0; u < kU; u++) {
		if (!ok[u])
			continue;
		for (u64 n = 0; old[u] != kEmpty; n++) {
			if (n >= kMaxProbe || n >= B.mask) {
				gatomic_or(B.error, 32u);
				break;
			}
			if ((u32)old[u] == off[u])
				*B.dups = 1u;
			h[u] = (h[u] + 1) & B.mask;
			old[u] = gatomic_cas(&B.slots[h[u]], kEmpty, val[u]);
		}
	}
}

} // namespace vh
