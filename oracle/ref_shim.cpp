// Test-only wrapper that exposes functions of the reference's own runtime
// (runtime.hpp / runtime.cpp of t1mm3/vldb_voila, compiled from
// /root/reference by oracle/Makefile into oracle/_ref/) with a C ABI, so the
// tests can pin the oracle's restatements against them.  Not shipped.
#include "runtime.hpp"

#include <cstdint>
#include <cstring>

extern "C" {

uint64_t ref_hash_i32(int32_t v) { return voila_hash<i32>::hash(v); }
uint64_t ref_hash_u32(uint32_t v) { return voila_hash<u32>::hash(v); }
uint64_t ref_hash_i64(int64_t v) { return voila_hash<i64>::hash(v); }
uint64_t ref_hash_u64(uint64_t v) { return voila_hash<u64>::hash(v); }
uint64_t ref_hash_u8(uint8_t v) { return voila_hash<u8>::hash(v); }
uint64_t ref_rehash_i32(uint64_t h, int32_t v) { return voila_hash<i32>::rehash(h, v); }
uint64_t ref_rehash_u32(uint64_t h, uint32_t v) { return voila_hash<u32>::rehash(h, v); }
uint64_t ref_rehash_i64(uint64_t h, int64_t v) { return voila_hash<i64>::rehash(h, v); }

// varchar(void* data, idx, max_len) over a column of (1 + max_len)-byte slots
static varchar slot_of(const uint8_t* col, size_t idx, size_t max_len)
{
	return varchar((void*)col, idx, max_len);
}

uint64_t ref_hash_varchar(const uint8_t* col, size_t idx, size_t max_len)
{
	return voila_hash<varchar>::hash(slot_of(col, idx, max_len));
}

uint64_t ref_rehash_varchar(uint64_t h, const uint8_t* col, size_t idx, size_t max_len)
{
	return voila_hash<varchar>::rehash(h, slot_of(col, idx, max_len));
}

int ref_varchar_contains(const uint8_t* col, size_t idx, size_t max_len, const char* needle, uint32_t nlen)
{
	varchar n;
	n.arr = needle;
	n.len = (uint16_t)nlen;
	return slot_of(col, idx, max_len).contains(n) ? 1 : 0;
}

int ref_varchar_eq(const uint8_t* col1, size_t idx1, size_t max_len1, const uint8_t* col2, size_t idx2,
	size_t max_len2)
{
	return slot_of(col1, idx1, max_len1) == slot_of(col2, idx2, max_len2) ? 1 : 0;
}

size_t ref_to_cstr_i128(uint64_t lo, uint64_t hi, char* out)
{
	i128 v = (i128)(((u128)hi << 64) | lo);
	char buf[voila_cast<i128>::good_buffer_size()];
	char* s = nullptr;
	size_t n = voila_cast<i128>::to_cstr(s, buf, sizeof(buf), v);
	memcpy(out, s, n);
	out[n] = 0;
	return n;
}

size_t ref_to_cstr_i64(int64_t v, char* out)
{
	char buf[voila_cast<i64>::good_buffer_size()];
	char* s = nullptr;
	size_t n = voila_cast<i64>::to_cstr(s, buf, sizeof(buf), v);
	memcpy(out, s, n);
	out[n] = 0;
	return n;
}

size_t ref_to_cstr_u64(uint64_t v, char* out)
{
	char buf[voila_cast<u64>::good_buffer_size()];
	char* s = nullptr;
	size_t n = voila_cast<u64>::to_cstr(s, buf, sizeof(buf), v);
	memcpy(out, s, n);
	out[n] = 0;
	return n;
}

int32_t ref_select_true(int32_t* sel, int32_t num, uint32_t* res, uint8_t* pred)
{
	return select_true(sel, num, res, pred);
}

int32_t ref_select_false(int32_t* sel, int32_t num, uint32_t* res, uint8_t* pred)
{
	return select_false(sel, num, res, pred);
}

int32_t ref_selunion(int32_t* result, int32_t* l, int32_t ln, int32_t* r, int32_t rn)
{
	return ComplexFuncs::selunion(result, l, ln, r, rn);
}

int32_t ref_avx512_hash_u32(int32_t* sel, int32_t num, uint64_t* res, uint32_t* col)
{
	return avx512_hash_u32(sel, num, res, col);
}

int32_t ref_bucket_lookup(int32_t* sel, int32_t num, uint64_t* res, const uint64_t* buckets,
	const uint64_t* index, uint64_t mask)
{
	return avx512_bucket_lookup(sel, num, res, buckets, index, mask);
}

int32_t ref_bucket_next(int32_t* sel, int32_t num, uint64_t* res, uint64_t* bucket,
	uint64_t next_offset)
{
	return avx512_bucket_next(sel, num, res, bucket, next_offset);
}

int32_t ref_check_u32(int32_t* sel, int32_t num, uint8_t* res, uint64_t* bucket,
	uint32_t* key, uint64_t offset)
{
	return avx512_check__u8__u32_col_u64_col_u64_col_u32_col(sel, num, res, bucket, key, offset);
}

}
