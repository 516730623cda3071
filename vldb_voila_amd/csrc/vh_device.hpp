// Device-side helpers shared by the gfx950 kernels: typed column loads,
// VOILA's hash, 64-lane wave primitives and exact 128-bit accumulation.
#pragma once

#ifndef __HIPCC_RTC__   // hipRTC provides these itself
#include <hip/hip_runtime.h>
#include <stdint.h>
#endif

#include "../../include/voila_hip.h"

namespace vh {

typedef unsigned long long u64;
typedef long long i64;
typedef unsigned int u32;

static constexpr int kWave = 64;

// Global memory through address-space-1 pointers.  Loads and atomics through
// generic pointers compile to FLAT instructions, which count against both
// the vector-memory and the LDS counters; while one is outstanding the
// compiler can only wait for all vector memory operations together, which
// serialises a loop that keeps loads in flight across iterations.  The
// atomics are relaxed and device-scoped, as HIP's atomicAdd etc. are.
#define VH_GLOBAL __attribute__((address_space(1)))

__device__ __forceinline__ u32 gatomic_or(u32* p, u32 v)
{
	return __hip_atomic_fetch_or((VH_GLOBAL u32*)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ u32 gatomic_and(u32* p, u32 v)
{
	return __hip_atomic_fetch_and((VH_GLOBAL u32*)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ u32 gatomic_add(u32* p, u32 v)
{
	return __hip_atomic_fetch_add((VH_GLOBAL u32*)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ u64 gatomic_add(u64* p, u64 v)
{
	return __hip_atomic_fetch_add((VH_GLOBAL u64*)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Compare-and-swap; returns the value found (== cmp when it swapped).
__device__ __forceinline__ u64 gatomic_cas(u64* p, u64 cmp, u64 v)
{
	__hip_atomic_compare_exchange_strong((VH_GLOBAL u64*)p, &cmp, v, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
		__HIP_MEMORY_SCOPE_AGENT);
	return cmp;
}

template <typename T>
__device__ __forceinline__ T gload(const T* p, i64 i)
{
	return ((const VH_GLOBAL T*)p)[i];
}

// voila_hash<T>::hash on the value converted to u64 (runtime.hpp:413-424).
__device__ __forceinline__ u64 voila_hash(u64 x)
{
	x ^= x >> 16;
	x *= 0x45d9f3bull;
	x ^= x >> 16;
	x *= 0x45d9f3bull;
	x ^= x >> 16;
	return x;
}

// voila_hash<T>::rehash (runtime.hpp:426-428).
__device__ __forceinline__ u64 voila_rehash(u64 h, u64 v)
{
	return voila_hash(v) ^ (h * 2654435761ull);
}

__device__ __forceinline__ int type_width(int t)
{
	switch (t) {
	case VH_U8: case VH_I8: return 1;
	case VH_U16: case VH_I16: return 2;
	case VH_U32: case VH_I32: return 4;
	default: return 8;
	}
}

// Canonical 64-bit value of element i of a typed column: signed types sign
// extend, unsigned types zero extend.
__device__ __forceinline__ i64 load1(const void* p, int t, i64 i)
{
	switch (t) {
	case VH_U8: return gload((const uint8_t*)p, i);
	case VH_I8: return gload((const int8_t*)p, i);
	case VH_U16: return gload((const uint16_t*)p, i);
	case VH_I16: return gload((const int16_t*)p, i);
	case VH_U32: return gload((const uint32_t*)p, i);
	case VH_I32: return gload((const int32_t*)p, i);
	default: return gload((const i64*)p, i);
	}
}

// load1 of N rows of one column: the type is dispatched once, so the N loads
// sit in one block and are issued together (load1 per row puts each load in
// a switch case of its own, and each waits for its value).
template <typename T, int N>
__device__ __forceinline__ void loadn_typed(const void* p, const i64 (&i)[N], i64 (&v)[N])
{
	T x[N];
#pragma unroll
	for (int u = 0; u < N; u++)
		x[u] = gload((const T*)p, i[u]);
#pragma unroll
	for (int u = 0; u < N; u++)
		v[u] = (i64)x[u];
}

template <int N>
__device__ __forceinline__ void loadn(const void* p, int t, const i64 (&i)[N], i64 (&v)[N])
{
	switch (t) {
	case VH_U8: loadn_typed<uint8_t, N>(p, i, v); break;
	case VH_I8: loadn_typed<int8_t, N>(p, i, v); break;
	case VH_U16: loadn_typed<uint16_t, N>(p, i, v); break;
	case VH_I16: loadn_typed<int16_t, N>(p, i, v); break;
	case VH_U32: loadn_typed<uint32_t, N>(p, i, v); break;
	case VH_I32: loadn_typed<int32_t, N>(p, i, v); break;
	default: loadn_typed<i64, N>(p, i, v); break;
	}
}

// Four consecutive elements [i, i+4) with one or two 16-byte loads
// (i % 4 == 0 and a 16-byte aligned column, checked on the host).
__device__ __forceinline__ void load4(const void* p, int t, i64 i, i64 v[4])
{
	switch (t) {
	case VH_U8: {
		uchar4 x = *(const uchar4*)((const uint8_t*)p + i);
		v[0] = x.x; v[1] = x.y; v[2] = x.z; v[3] = x.w;
		break;
	}
	case VH_I8: {
		char4 x = *(const char4*)((const int8_t*)p + i);
		v[0] = x.x; v[1] = x.y; v[2] = x.z; v[3] = x.w;
		break;
	}
	case VH_U16: {
		ushort4 x = *(const ushort4*)((const uint16_t*)p + i);
		v[0] = x.x; v[1] = x.y; v[2] = x.z; v[3] = x.w;
		break;
	}
	case VH_I16: {
		short4 x = *(const short4*)((const int16_t*)p + i);
		v[0] = x.x; v[1] = x.y; v[2] = x.z; v[3] = x.w;
		break;
	}
	case VH_U32: {
		uint4 x = *(const uint4*)((const uint32_t*)p + i);
		v[0] = x.x; v[1] = x.y; v[2] = x.z; v[3] = x.w;
		break;
	}
	case VH_I32: {
		int4 x = *(const int4*)((const int32_t*)p + i);
		v[0] = x.x; v[1] = x.y; v[2] = x.z; v[3] = x.w;
		break;
	}
	default: {
		const longlong2* q = (const longlong2*)((const i64*)p + i);
		longlong2 a = q[0], b = q[1];
		v[0] = a.x; v[1] = a.y; v[2] = b.x; v[3] = b.y;
		break;
	}
	}
}

// Predicate value `a cmp b` on canonical values; u64 compares unsigned.
__device__ __forceinline__ bool compare(int cmp, int t, i64 a, i64 b)
{
	if (t == VH_U64) {
		u64 x = (u64)a, y = (u64)b;
		switch (cmp) {
		case VH_LT: return x < y;
		case VH_LE: return x <= y;
		case VH_GT: return x > y;
		case VH_GE: return x >= y;
		case VH_EQ: return x == y;
		default: return x != y;
		}
	}
	switch (cmp) {
	case VH_LT: return a < b;
	case VH_LE: return a <= b;
	case VH_GT: return a > b;
	case VH_GE: return a >= b;
	case VH_EQ: return a == b;
	default: return a != b;
	}
}

// Exact 128-bit accumulation into (lo: u64, hi: i64) words in global
// memory: the low word is added atomically, its carry and the sign extension
// of the addend go to the high word.  Integer addition commutes, so the
// result is independent of the order in which blocks merge.
__device__ __forceinline__ void atomic_add_i128(u64* lo, i64* hi, u64 add_lo, i64 add_hi)
{
	u64 old = gatomic_add(lo, add_lo);
	i64 carry = (old + add_lo) < old ? 1 : 0;
	i64 h = add_hi + carry;
	if (h)
		gatomic_add((u64*)hi, (u64)h);
}

__device__ __forceinline__ u64 shfl_xor_u64(u64 v, int m)
{
	u32 lo = (u32)v, hi = (u32)(v >> 32);
	lo = __shfl_xor(lo, m, kWave);
	hi = __shfl_xor(hi, m, kWave);
	return ((u64)hi << 32) | lo;
}

// Exact wave-wide sum of (lo, hi) pairs: hi collects carries.
__device__ __forceinline__ void wave_sum_i128(u64& lo, i64& hi)
{
#pragma unroll
	for (int m = 32; m >= 1; m >>= 1) {
		u64 olo = shfl_xor_u64(lo, m);
		i64 ohi = (i64)shfl_xor_u64((u64)hi, m);
		u64 nlo = lo + olo;
		hi = hi + ohi + (nlo < lo ? 1 : 0);
		lo = nlo;
	}
}

__device__ __forceinline__ u64 wave_sum_u64(u64 v)
{
#pragma unroll
	for (int m = 32; m >= 1; m >>= 1)
		v += shfl_xor_u64(v, m);
	return v;
}

} // namespace vh
