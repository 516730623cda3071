// Host-side declarations of the statement-level kernel launchers
// (vh_primitives.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace vh {

typedef unsigned long long u64;
typedef long long i64;
typedef unsigned int u32;

static constexpr i64 kSelTileRows = 4096;   // positions per selection tile
i64 sel_tiles(i64 inum);
hipError_t launch_select(const int32_t* sel, i64 inum, const uint8_t* pred, int want, int32_t* res,
	u32* tile_scratch, int32_t* d_count, hipStream_t s);
hipError_t launch_cmp(int cmp, int t, const int32_t* sel, i64 inum, uint8_t* res, const void* a,
	const void* b, i64 c, hipStream_t s);
hipError_t launch_arith(int op, int rt, const int32_t* sel, i64 inum, void* res, int ta, const void* a,
	int tb, const void* b, i64 cb, hipStream_t s);
hipError_t launch_hash(int t, const int32_t* sel, i64 inum, u64* res, const u64* h, const void* col,
	hipStream_t s);
hipError_t launch_bucket_lookup(const u64* heads, u64 mask, const int32_t* sel, i64 inum, u64* res,
	const u64* hash, hipStream_t s);
hipError_t launch_bucket_next(const u64* next, u64 rows, const int32_t* sel, i64 inum, u64* res,
	const u64* bucket, hipStream_t s);
hipError_t launch_check(const i64* colv, u64 rows, const int32_t* sel, i64 inum, uint8_t* res,
	const u64* bucket, int kt, const void* key, hipStream_t s);
hipError_t launch_gather(const i64* colv, u64 rows, const int32_t* sel, i64 inum, void* res, int rt,
	const u64* bucket, hipStream_t s);
hipError_t launch_scatter(i64* colv, u64 rows, const int32_t* sel, i64 inum, const u64* bucket, int vt,
	const void* val, hipStream_t s);
hipError_t launch_aggr(u64* col2, u64 rows, const int32_t* sel, i64 inum, const u64* bucket, int vt,
	const void* val, hipStream_t s);
hipError_t launch_bucket_insert(u64* heads, u64 mask, u64* next, u64* hashcol, u64 max_rows, u64* counters,
	u32* overflow, const int32_t* sel, i64 inum, u64* res, const u64* hash, hipStream_t s);
hipError_t launch_write(i64* dst, const int32_t* sel, i64 inum, int vt, const void* val, hipStream_t s);
hipError_t launch_index_build(u64* heads, u64 mask, u64* next, const u64* hashcol, u64 rows, hipStream_t s);

} // namespace vh
