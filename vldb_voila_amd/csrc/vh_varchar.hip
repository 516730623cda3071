// Statement-level gfx950 kernels over varchar columns: vec_hash / vec_rehash
// (MurmurHash64A), vec_contains and vec_eq / vec_ne.
//
// Column layout: one (1 + max_len)-byte slot per row, the length byte first,
// then the bytes, unused tail bytes zero.  This is the reference's binary
// column of Char<n> / Varchar<n> (a uint8_t length indicator then `char
// value[n]`, db-engine-paradigms include/common/runtime/Types.hpp:125-130,
// :243-250, written and read back whole by Import.cpp:165-190) that
// varchar(void* data, idx, max_len) addresses (runtime.hpp:192-205).
//
// A workgroup takes a tile of consecutive positions.  For a dense range
// (sel == nullptr) the tile's slots are one contiguous byte range, which the
// workgroup first copies into LDS with coalesced dword loads; each lane then
// reads its row from LDS.  With a selection vector, lanes read their own
// slots from global memory.
#include "vh_device.hpp"
#include "vh_varchar.hpp"

namespace vh {

static constexpr int kVcThreads = 256;
static constexpr int kVcTileBytes = 32768;   // LDS staging per workgroup

struct VcArgs {
	const int32_t* sel;
	i64 inum;
	const uint8_t* col;         // varchar column: (1 + max_len)-byte slots
	int slot;
	int tile_rows;         // rows staged per tile (tile_rows * slot <= kVcTileBytes)
	const u64* h;          // rehash: previous hashes (nullptr: hash, seed 0)
	void* res;
	int op;                // VC_HASH, VC_CONTAINS, VC_EQ, VC_NE
	const uint8_t* col2;        // VC_EQ / VC_NE against a column (nullptr: the constant)
	int slot2;
	u32 clen;              // needle / constant length
	uint8_t cbytes[kVcMaxConst];
};

// MurmurHash64A (voila_hash<varchar>::hash, runtime.hpp:434-481): 8-byte
// little-endian blocks, then the len & 7 tail bytes, then the final mix.
__device__ __forceinline__ u64 murmur64a(const uint8_t* p, u32 len, u64 seed)
{
	const u64 m = 0xc6a4a7935bd1e995ull;
	const int r = 47;
	u64 h = seed ^ ((u64)len * m);
	const u32 nb = len >> 3;
	for (u32 b = 0; b < nb; b++) {
		u64 k = 0;
#pragma unroll
		for (int j = 0; j < 8; j++)
			k |= (u64)p[8 * b + j] << (8 * j);
		k *= m;
		k ^= k >> r;
		k *= m;
		h ^= k;
		h *= m;
	}
	const u32 rem = len & 7;
	if (rem) {
		for (u32 j = 0; j < rem; j++)
			h ^= (u64)p[8 * nb + j] << (8 * j);
		h *= m;
	}
	h ^= h >> r;
	h *= m;
	h ^= h >> r;
	return h;
}

// varchar::contains: memmem(arr, len, needle, nlen) != NULL (runtime.hpp:206-208);
// an empty needle is found in every string.
__device__ __forceinline__ bool vc_contains(const uint8_t* p, u32 len, const uint8_t* needle, u32 nlen)
{
	if (nlen > len)
		return false;
	for (u32 s = 0; s + nlen <= len; s++) {
		u32 j = 0;
		while (j < nlen && p[s + j] == needle[j])
			j++;
		if (j == nlen)
			return true;
	}
	return false;
}

// varchar::operator== (runtime.hpp:154-156): equal lengths and
// strncmp(a, b, len) == 0, which stops at a NUL both strings share.
__device__ __forceinline__ bool vc_equal(const uint8_t* a, u32 alen, const uint8_t* b, u32 blen)
{
	if (alen != blen)
		return false;
	for (u32 j = 0; j < alen; j++) {
		if (a[j] != b[j])
			return false;
		if (a[j] == 0)
			return true;
	}
	return true;
}

// ---- the same three operations on a row staged in LDS, four bytes at a time

// Bytes i .. i + 3 of the LDS tile (any alignment): two aligned dwords and a
// byte funnel shift.
__device__ __forceinline__ u32 lds4(const u32* w, u32 i)
{
	return __builtin_amdgcn_alignbyte(w[(i >> 2) + 1], w[i >> 2], i & 3);
}

__device__ __forceinline__ u32 low_bytes(u32 n)   // mask of the low n bytes, n <= 4
{
	return n >= 4 ? 0xFFFFFFFFu : (1u << (8 * n)) - 1u;
}

// 0x80 in every byte of x that is zero (exact: no carry between bytes)
__device__ __forceinline__ u32 zero_bytes(u32 x)
{
	return ~(((x & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | x | 0x7F7F7F7Fu);
}

__device__ __forceinline__ u64 murmur64a_lds(const u32* w, u32 i, u32 len, u64 seed)
{
	const u64 m = 0xc6a4a7935bd1e995ull;
	const int r = 47;
	u64 h = seed ^ ((u64)len * m);
	const u32 nb = len >> 3;
	for (u32 b = 0; b < nb; b++) {
		u64 k = (u64)lds4(w, i + 8 * b) | ((u64)lds4(w, i + 8 * b + 4) << 32);
		k *= m;
		k ^= k >> r;
		k *= m;
		h ^= k;
		h *= m;
	}
	const u32 rem = len & 7;
	if (rem) {
		const u32 t = i + 8 * nb;
		u64 tail = lds4(w, t) & low_bytes(rem);
		if (rem > 4)
			tail |= (u64)(lds4(w, t + 4) & low_bytes(rem - 4)) << 32;
		h ^= tail;
		h *= m;
	}
	h ^= h >> r;
	h *= m;
	h ^= h >> r;
	return h;
}

// memmem: candidate starts are the bytes equal to the needle's first byte
// (found four at a time), each then compared four bytes at a time.
__device__ __forceinline__ bool contains_lds(const u32* w, u32 i, u32 len, const u32* nw, u32 nlen)
{
	if (nlen > len)
		return false;
	if (nlen == 0)
		return true;
	const u32 last = len - nlen;                        // the last start
	const u32 first = (lds4(nw, 0) & 0xFFu) * 0x01010101u;
	for (u32 s0 = 0; s0 <= last; s0 += 8) {
		const u32 w0 = lds4(w, i + s0);
		const u32 w1 = lds4(w, i + s0 + 4);
		u64 t = (u64)zero_bytes(w0 ^ first) | ((u64)zero_bytes(w1 ^ first) << 32);
		if (last - s0 < 7)
			t &= (1ull << (8 * (last - s0 + 1))) - 1ull;   // starts s0 .. last only
		while (t) {
			const u32 s = s0 + ((u32)__builtin_ctzll(t) >> 3);
			bool ok = true;
			for (u32 j = 1; j < nlen && ok; j += 4)
				ok = ((lds4(w, i + s + j) ^ lds4(nw, j)) & low_bytes(nlen - j)) == 0;
			if (ok)
				return true;
			t &= t - 1;
		}
	}
	return false;
}

// varchar::operator== against the constant: equal lengths, then bytes up to
// the first difference or a NUL both share.
__device__ __forceinline__ bool equal_lds(const u32* w, u32 i, u32 len, const u32* cw, u32 clen)
{
	if (len != clen)
		return false;
	for (u32 j = 0; j < len; j += 4) {
		const u32 a = lds4(w, i + j);
		const u32 hb = low_bytes(len - j) & 0x80808080u;
		const u32 ne = ~zero_bytes(a ^ lds4(cw, j)) & hb;   // bytes that differ
		const u32 nul = zero_bytes(a) & hb;                  // NUL bytes of the row
		if (ne | nul) {
			const u32 fne = ne ? (u32)__builtin_ctz(ne) : 32u;
			const u32 fnul = nul ? (u32)__builtin_ctz(nul) : 32u;
			return fnul < fne;   // a NUL both share, before any difference: equal
		}
	}
	return true;
}

// Copy slots [r0, r0 + cnt) of a dense column of `total` bytes into LDS in
// 16-byte chunks, each thread's (up to) eight loads in flight together; a
// chunk that reaches past either end of the column is read byte by byte.
// Returns the LDS offset of row r0's slot.
__device__ __forceinline__ u32 stage_slots(const uint8_t* col, u64 total, i64 r0, int cnt, int slot, u32* lds)
{
	const uintptr_t base = (uintptr_t)col;
	const uintptr_t b0 = base + (u64)r0 * (u64)slot;
	const uintptr_t b1 = b0 + (u64)cnt * (u64)slot;
	const uintptr_t a0 = b0 & ~(uintptr_t)15;
	const u32 chunks = (u32)((((b1 + 15) & ~(uintptr_t)15) - a0) / 16);
	uint4* dst = (uint4*)lds;
	for (u32 c0 = 0; c0 < chunks; c0 += 8 * kVcThreads) {
		uint4 v[8];
#pragma unroll
		for (int u = 0; u < 8; u++) {
			const u32 c = c0 + u * kVcThreads + threadIdx.x;
			const uintptr_t a = a0 + 16 * (uintptr_t)c;
			v[u] = make_uint4(0u, 0u, 0u, 0u);
			if (c < chunks && a >= base && a + 16 <= base + total)
				v[u] = *(const uint4*)a;
		}
#pragma unroll
		for (int u = 0; u < 8; u++) {
			const u32 c = c0 + u * kVcThreads + threadIdx.x;
			if (c >= chunks)
				continue;
			const uintptr_t a = a0 + 16 * (uintptr_t)c;
			if (!(a >= base && a + 16 <= base + total)) {
				u32 wv[4] = {0u, 0u, 0u, 0u};
				for (int b = 0; b < 16; b++)
					if (a + b >= base && a + b < base + total)
						wv[b >> 2] |= (u32) * (const uint8_t*)(a + b) << (8 * (b & 3));
				v[u] = make_uint4(wv[0], wv[1], wv[2], wv[3]);
			}
			dst[c] = v[u];
		}
	}
	__syncthreads();
	return (u32)(b0 - a0);
}

__global__ __launch_bounds__(kVcThreads) void varchar_kernel(VcArgs A)
{
	extern __shared__ uint4 vc_lds16[];        // 16-byte aligned: stage_slots stores uint4
	u32* vc_lds = (u32*)vc_lds16;
	__shared__ u32 cst_words[kVcMaxConst / 4 + 2];    // + 2: lds4 reads a dword past the last byte
	uint8_t* cst = (uint8_t*)cst_words;
	for (u32 j = threadIdx.x; j < sizeof(cst_words); j += blockDim.x)
		cst[j] = j < A.clen ? A.cbytes[j] : 0;
	__syncthreads();
	const int R = A.tile_rows;
	const u32 maxlen = (u32)A.slot - 1;
	for (i64 t0 = (i64)blockIdx.x * R; t0 < A.inum; t0 += (i64)gridDim.x * R) {
		const int cnt = (int)(A.inum - t0 < R ? A.inum - t0 : R);
		u32 off = 0;
		if (!A.sel)
			off = stage_slots(A.col, (u64)A.inum * (u64)A.slot, t0, cnt, A.slot, vc_lds);
		if ((int)threadIdx.x < cnt) {
			const i64 k = t0 + threadIdx.x;
			const i64 i = A.sel ? (i64)A.sel[k] : k;
			const uint8_t* p = A.sel ? A.col + (u64)i * (u64)A.slot
			                    : (const uint8_t*)vc_lds + off + (u32)threadIdx.x * (u32)A.slot;
			// the length byte, bounded by the slot (well-formed columns never exceed it)
			const u32 len = p[0] < maxlen ? p[0] : maxlen;
			const uint8_t* s = p + 1;
			const u32 li = off + (u32)threadIdx.x * (u32)A.slot + 1;   // the row's bytes in the tile
			if (!A.sel && A.op == VC_HASH) {
				((u64*)A.res)[i] = murmur64a_lds(vc_lds, li, len, A.h ? A.h[i] : 0ull);
			} else if (!A.sel && A.op == VC_CONTAINS) {
				((uint8_t*)A.res)[i] = contains_lds(vc_lds, li, len, cst_words, A.clen) ? 1 : 0;
			} else if (!A.sel && !A.col2) {
				const bool eq = equal_lds(vc_lds, li, len, cst_words, A.clen);
				((uint8_t*)A.res)[i] = (eq == (A.op == VC_EQ)) ? 1 : 0;
			} else if (A.op == VC_HASH) {
				((u64*)A.res)[i] = murmur64a(s, len, A.h ? A.h[i] : 0ull);
			} else if (A.op == VC_CONTAINS) {
				((uint8_t*)A.res)[i] = vc_contains(s, len, cst, A.clen) ? 1 : 0;
			} else {
				bool eq;
				if (A.col2) {
					const uint8_t* q = A.col2 + (u64)i * (u64)A.slot2;
					const u32 m2 = (u32)A.slot2 - 1;
					eq = vc_equal(s, len, q + 1, q[0] < m2 ? q[0] : m2);
				} else {
					eq = vc_equal(s, len, cst, A.clen);
				}
				((uint8_t*)A.res)[i] = (eq == (A.op == VC_EQ)) ? 1 : 0;
			}
		}
		__syncthreads();   // the LDS tile is reused by the next iteration
	}
}

hipError_t launch_varchar(int op, const int32_t* sel, i64 inum, void* res, const uint8_t* col, int max_len,
	const u64* h, const uint8_t* col2, int max_len2, const uint8_t* cbytes, u32 clen, hipStream_t s)
{
	if (inum <= 0)
		return hipSuccess;
	VcArgs A = {};
	A.sel = sel;
	A.inum = inum;
	A.col = col;
	A.slot = max_len + 1;
	A.tile_rows = kVcTileBytes / A.slot < kVcThreads ? kVcTileBytes / A.slot : kVcThreads;
	A.h = h;
	A.res = res;
	A.op = op;
	A.col2 = col2;
	A.slot2 = max_len2 + 1;
	A.clen = clen;
	for (u32 j = 0; j < clen; j++)
		A.cbytes[j] = cbytes[j];
	i64 grid = (inum + A.tile_rows - 1) / A.tile_rows;
	if (grid > 8192)
		grid = 8192;
	// staging: the tile's slots, up to 15 bytes either side of 16-byte
	// alignment, and the dword past the last byte that lds4 reads
	const size_t lds = sel ? 0 : ((size_t)A.tile_rows * A.slot + 48 + 15) / 16 * 16;
	hipLaunchKernelGGL(varchar_kernel, dim3((unsigned)grid), dim3(kVcThreads), lds, s, A);
	return hipGetLastError();
}

} // namespace vh
