// extern "C" statement-level entry points: the vec_* primitive family over
// whole device columns, and the device chained hash table they address.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <utility>
#include <vector>

#include "../../include/voila_hip.h"
#include "vh_engine.hpp"
#include "vh_kernels.hpp"
#include "vh_prims.hpp"
#include "vh_varchar.hpp"

using namespace vh;

#define HIP_CALL(x)                                                                  \
	do {                                                                             \
		hipError_t _e = (x);                                                         \
		if (_e != hipSuccess)                                                        \
			return fail(VH_ERR_HIP, "%s: %s", #x, hipGetErrorString(_e));           \
	} while (0)

struct vh_table {
	int ncols = 0;
	std::vector<int> wide;          // 1: 128-bit column (lo, hi words)
	std::vector<u64*> cols;
	u64* heads = nullptr;
	u64 mask = 0;
	u64* next = nullptr;
	u64* hash = nullptr;
	u64* counters = nullptr;        // [0] rows, [1] first row of the running insert
	u32* overflow = nullptr;
	u64 max_rows = 0;
	int64_t host_rows = 0;          // rows appended by vh_vec_write (write_pos)
	std::vector<void*> allocs;
	~vh_table()
	{
		for (void* p : allocs)
			(void)hipFree(p);
	}
	int alloc(void** p, size_t n)
	{
		hipError_t e = hipMalloc(p, std::max<size_t>(n, 16));
		if (e != hipSuccess)
			return fail(VH_ERR_NOMEM, "hipMalloc(%zu): %s", n, hipGetErrorString(e));
		allocs.push_back(*p);
		return VH_OK;
	}
};

// calc_num_buckets (runtime_struct.cpp:790-814)
static u64 num_buckets(u64 count)
{
	u64 n = std::max<u64>({count, 2048, 2048});
	int exp = 64 - __builtin_clzll(n);
	if ((double)(1ull << exp) < (double)n / 0.7)
		exp++;
	return 1ull << exp;
}

// Tile-count scratch of the selections (vh_vec_seltrue / vh_vec_selfalse),
// one buffer per (device, stream) of the calling thread, so selections
// enqueued on different streams never share one.  A buffer that must grow is
// not freed: work already enqueued (or captured in a graph) may still read
// it, so it is retired and kept; growth at least doubles, so the retired
// buffers of a stream hold at most as much as its current one.
struct SelScratch {
	u32* tiles = nullptr;
	int64_t cap = 0;
};
static thread_local std::map<std::pair<int, hipStream_t>, SelScratch> g_sel_scratch;
static thread_local std::vector<void*> g_sel_retired;

static int sel_scratch(int64_t inum, hipStream_t stream, u32** out)
{
	const int64_t need = sel_tiles(inum) + 1;
	int dev = 0;
	HIP_CALL(hipGetDevice(&dev));
	SelScratch& S = g_sel_scratch[std::make_pair(dev, stream)];
	if (need > S.cap) {
		const int64_t cap = std::max<int64_t>(need, 2 * S.cap);
		u32* p = nullptr;
		hipError_t e = hipMalloc((void**)&p, cap * sizeof(u32));
		if (e != hipSuccess)
			return fail(VH_ERR_NOMEM, "selection scratch: %s", hipGetErrorString(e));
		if (S.tiles)
			g_sel_retired.push_back(S.tiles);
		S.tiles = p;
		S.cap = cap;
	}
	*out = S.tiles;
	return VH_OK;
}

// Tile-count scratch of vh_vec_seltrue_noalloc: a static device array, so the
// call allocates nothing.  Each device holds its own copy (at its own
// address) until the process ends.  In host code `scratch` only names the
// symbol; its device address comes from hipGetSymbolAddress, looked up once
// per device (with that device current) and kept in g_scratch_ptr.
static constexpr int64_t kStaticTiles = 131072;   // 512 KiB of u32 tile counts
static constexpr int kMaxDevices = 64;
static __device__ __align__(256) u32 scratch[kStaticTiles];
static std::mutex g_scratch_mu;
static u32* g_scratch_ptr[kMaxDevices];

static int static_scratch(u32** out)
{
	int dev = 0;
	HIP_CALL(hipGetDevice(&dev));
	if (dev < 0 || dev >= kMaxDevices)
		return fail(VH_ERR_INVALID, "device %d: at most %d devices", dev, kMaxDevices);
	std::lock_guard<std::mutex> lock(g_scratch_mu);
	if (!g_scratch_ptr[dev]) {
		void* p = nullptr;
		hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
		if (e != hipSuccess || !p)
			return fail(VH_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
		g_scratch_ptr[dev] = (u32*)p;
	}
	*out = g_scratch_ptr[dev];
	return VH_OK;
}

static int check_table(vh_table* t, int col)
{
	if (!t)
		return fail(VH_ERR_INVALID, "null table");
	if (col < 0 || col >= t->ncols)
		return fail(VH_ERR_INVALID, "column %d of %d", col, t->ncols);
	return VH_OK;
}

#define LAUNCH(x)                                                                   \
	do {                                                                            \
		hipError_t _e = (x);                                                        \
		if (_e != hipSuccess)                                                       \
			return fail(VH_ERR_HIP, "%s: %s", #x, hipGetErrorString(_e));          \
		return VH_OK;                                                               \
	} while (0)

static bool valid_type(int t) { return t >= VH_U8 && t <= VH_I64; }

extern "C" {

int vh_table_new(vh_table** out, int32_t ncols, const int32_t* wide, int64_t max_rows)
{
	if (!out || ncols < 0 || ncols > 64 || max_rows < 0)
		return fail(VH_ERR_INVALID, "vh_table_new arguments");
	vh_table* t = new (std::nothrow) vh_table();
	if (!t)
		return fail(VH_ERR_NOMEM, "host allocation");
	t->ncols = ncols;
	t->max_rows = (u64)max_rows;
	const u64 cap = num_buckets((u64)max_rows);
	t->mask = cap - 1;
	int rc = t->alloc((void**)&t->heads, cap * 8);
	if (!rc) rc = t->alloc((void**)&t->next, (size_t)max_rows * 8);
	if (!rc) rc = t->alloc((void**)&t->hash, (size_t)max_rows * 8);
	if (!rc) rc = t->alloc((void**)&t->counters, 16);
	if (!rc) rc = t->alloc((void**)&t->overflow, 16);
	for (int c = 0; c < ncols && !rc; c++) {
		const int w = wide ? wide[c] != 0 : 0;
		t->wide.push_back(w);
		u64* p = nullptr;
		rc = t->alloc((void**)&p, (size_t)max_rows * 8 * (w ? 2 : 1));
		t->cols.push_back(p);
	}
	if (rc) {
		delete t;
		return rc;
	}
	*out = t;
	return vh_table_reset(t, nullptr);
}

int vh_table_free(vh_table* t)
{
	delete t;
	return VH_OK;
}

// ITable::reset (runtime_struct.cpp:1160-1170): empty rows, empty index.
int vh_table_reset(vh_table* t, void* stream)
{
	if (!t)
		return fail(VH_ERR_INVALID, "null table");
	hipStream_t s = (hipStream_t)stream;
	HIP_CALL(launch_fill(t->heads, t->mask + 1, 0, s));
	HIP_CALL(launch_fill(t->counters, 2, 0, s));
	HIP_CALL(launch_fill((u64*)t->overflow, 2, 0, s));
	for (int c = 0; c < t->ncols; c++)
		HIP_CALL(launch_fill(t->cols[c], t->max_rows * (t->wide[c] ? 2 : 1), 0, s));
	t->host_rows = 0;
	return VH_OK;
}

int vh_table_rows(vh_table* t, void* stream, int64_t* rows)
{
	if (!t || !rows)
		return fail(VH_ERR_INVALID, "null argument");
	HIP_CALL(hipStreamSynchronize((hipStream_t)stream));
	u64 c[2];
	u32 of[4];
	HIP_CALL(hipMemcpy(c, t->counters, 16, hipMemcpyDeviceToHost));
	HIP_CALL(hipMemcpy(of, t->overflow, 16, hipMemcpyDeviceToHost));
	if (of[0])
		return fail(VH_ERR_CAPACITY, "table full (%llu rows)", (unsigned long long)t->max_rows);
	*rows = (int64_t)std::max<u64>(c[0], (u64)t->host_rows);
	return VH_OK;
}

int vh_table_column(vh_table* t, int32_t col, void** dev_ptr)
{
	int rc = check_table(t, col);
	if (rc)
		return rc;
	*dev_ptr = t->cols[col];
	return VH_OK;
}

// --- selection ---------------------------------------------------------------

int vh_vec_seltrue(const int32_t* sel, int64_t inum, int32_t* res, const uint8_t* pred, int32_t* d_onum,
	void* stream)
{
	if (inum < 0 || inum > INT32_MAX || !d_onum || (inum && (!res || !pred)))
		return fail(VH_ERR_INVALID, "vh_vec_seltrue arguments");
	u32* tiles;
	int rc = sel_scratch(inum, (hipStream_t)stream, &tiles);
	if (rc)
		return rc;
	LAUNCH(launch_select(sel, inum, pred, 1, res, tiles, d_onum, (hipStream_t)stream));
}

int vh_vec_seltrue_noalloc(const int32_t* sel, int64_t inum, int32_t* res, const uint8_t* pred,
	int32_t* d_onum, void* stream)
{
	if (inum < 0 || inum > INT32_MAX || !d_onum || (inum && (!res || !pred)))
		return fail(VH_ERR_INVALID, "vh_vec_seltrue_noalloc arguments");
	if (sel_tiles(inum) + 1 > kStaticTiles)
		return fail(VH_ERR_CAPACITY, "vh_vec_seltrue_noalloc: %lld positions exceed its static scratch "
			"(at most %lld)", (long long)inum, (long long)((kStaticTiles - 1) * kSelTileRows));
	u32* tiles;
	int rc = static_scratch(&tiles);
	if (rc)
		return rc;
	// every tile count the scan reads is written first by the count pass, so
	// the array needs no clearing between calls
	LAUNCH(launch_select(sel, inum, pred, 1, res, tiles, d_onum, (hipStream_t)stream));
}

int vh_vec_selfalse(const int32_t* sel, int64_t inum, int32_t* res, const uint8_t* pred, int32_t* d_onum,
	void* stream)
{
	if (inum < 0 || inum > INT32_MAX || !d_onum || (inum && (!res || !pred)))
		return fail(VH_ERR_INVALID, "vh_vec_selfalse arguments");
	u32* tiles;
	int rc = sel_scratch(inum, (hipStream_t)stream, &tiles);
	if (rc)
		return rc;
	LAUNCH(launch_select(sel, inum, pred, 0, res, tiles, d_onum, (hipStream_t)stream));
}

// --- maps ----------------------------------------------------------------------

int vh_vec_cmp(int32_t cmp, int32_t type, const int32_t* sel, int64_t inum, uint8_t* res,
	const void* col1, const void* col2, int64_t constant, void* stream)
{
	if (cmp < VH_LT || cmp > VH_NE || !valid_type(type) || inum < 0)
		return fail(VH_ERR_INVALID, "vh_vec_cmp arguments");
	LAUNCH(launch_cmp(cmp, type, sel, inum, res, col1, col2, constant, (hipStream_t)stream));
}

int vh_vec_arith(int32_t op, int32_t res_type, const int32_t* sel, int64_t inum, void* res, int32_t t1,
	const void* col1, int32_t t2, const void* col2, int64_t constant, void* stream)
{
	if (op < VH_E_ADD || op > VH_E_MUL || !valid_type(res_type) || !valid_type(t1) ||
		(col2 && !valid_type(t2)) || inum < 0)
		return fail(VH_ERR_INVALID, "vh_vec_arith arguments");
	LAUNCH(launch_arith(op, res_type, sel, inum, res, t1, col1, t2, col2, constant, (hipStream_t)stream));
}

int vh_vec_hash(int32_t type, const int32_t* sel, int64_t inum, uint64_t* res, const void* col, void* stream)
{
	if (!valid_type(type) || inum < 0)
		return fail(VH_ERR_INVALID, "vh_vec_hash arguments");
	LAUNCH(launch_hash(type, sel, inum, (u64*)res, nullptr, col, (hipStream_t)stream));
}

int vh_vec_rehash(int32_t type, const int32_t* sel, int64_t inum, uint64_t* res, const uint64_t* hash,
	const void* col, void* stream)
{
	if (!valid_type(type) || inum < 0 || !hash)
		return fail(VH_ERR_INVALID, "vh_vec_rehash arguments");
	LAUNCH(launch_hash(type, sel, inum, (u64*)res, (const u64*)hash, col, (hipStream_t)stream));
}

// --- varchar statements --------------------------------------------------------

static bool valid_max_len(int32_t n) { return n >= 1 && n <= kVcMaxLen; }

int vh_vec_hash_varchar(const int32_t* sel, int64_t inum, uint64_t* res, const uint8_t* col, int32_t max_len,
	void* stream)
{
	if (!valid_max_len(max_len) || inum < 0 || (inum && (!col || !res)))
		return fail(VH_ERR_INVALID, "vh_vec_hash_varchar arguments");
	LAUNCH(launch_varchar(VC_HASH, sel, inum, res, col, max_len, nullptr, nullptr, 0, nullptr, 0,
		(hipStream_t)stream));
}

int vh_vec_rehash_varchar(const int32_t* sel, int64_t inum, uint64_t* res, const uint64_t* hash,
	const uint8_t* col, int32_t max_len, void* stream)
{
	if (!valid_max_len(max_len) || inum < 0 || !hash || (inum && (!col || !res)))
		return fail(VH_ERR_INVALID, "vh_vec_rehash_varchar arguments");
	LAUNCH(launch_varchar(VC_HASH, sel, inum, res, col, max_len, (const u64*)hash, nullptr, 0, nullptr, 0,
		(hipStream_t)stream));
}

int vh_vec_contains(const int32_t* sel, int64_t inum, uint8_t* res, const uint8_t* col, int32_t max_len,
	const char* needle, int32_t needle_len, void* stream)
{
	if (!valid_max_len(max_len) || inum < 0 || needle_len < 0 || needle_len > kVcMaxLen ||
		(needle_len && !needle) || (inum && (!col || !res)))
		return fail(VH_ERR_INVALID, "vh_vec_contains arguments");
	LAUNCH(launch_varchar(VC_CONTAINS, sel, inum, res, col, max_len, nullptr, nullptr, 0, (const uint8_t*)needle,
		(unsigned)needle_len, (hipStream_t)stream));
}

int vh_vec_cmp_varchar(int32_t cmp, const int32_t* sel, int64_t inum, uint8_t* res, const uint8_t* col1,
	int32_t max_len1, const uint8_t* col2, int32_t max_len2, const char* constant, int32_t constant_len,
	void* stream)
{
	if ((cmp != VH_EQ && cmp != VH_NE) || !valid_max_len(max_len1) || inum < 0 || (inum && (!col1 || !res)))
		return fail(VH_ERR_INVALID, "vh_vec_cmp_varchar arguments");
	if (col2 ? !valid_max_len(max_len2)
	         : (constant_len < 0 || constant_len > kVcMaxLen || (constant_len && !constant)))
		return fail(VH_ERR_INVALID, "vh_vec_cmp_varchar operand");
	LAUNCH(launch_varchar(cmp == VH_EQ ? VC_EQ : VC_NE, sel, inum, res, col1, max_len1, nullptr, col2, max_len2,
		col2 ? nullptr : (const uint8_t*)constant, col2 ? 0u : (unsigned)constant_len, (hipStream_t)stream));
}

// --- table statements ----------------------------------------------------------

int vh_vec_bucket_lookup(vh_table* t, const int32_t* sel, int64_t inum, uint64_t* res, const uint64_t* hash,
	void* stream)
{
	if (!t || inum < 0)
		return fail(VH_ERR_INVALID, "vh_vec_bucket_lookup arguments");
	LAUNCH(launch_bucket_lookup(t->heads, t->mask, sel, inum, (u64*)res, (const u64*)hash, (hipStream_t)stream));
}

int vh_vec_bucket_next(vh_table* t, const int32_t* sel, int64_t inum, uint64_t* res, const uint64_t* bucket,
	void* stream)
{
	if (!t || inum < 0)
		return fail(VH_ERR_INVALID, "vh_vec_bucket_next arguments");
	LAUNCH(launch_bucket_next(t->next, t->max_rows, sel, inum, (u64*)res, (const u64*)bucket,
		(hipStream_t)stream));
}

int vh_vec_bucket_insert(vh_table* t, const int32_t* sel, int64_t inum, uint64_t* res, const uint64_t* hash,
	void* stream)
{
	if (!t || inum < 0)
		return fail(VH_ERR_INVALID, "vh_vec_bucket_insert arguments");
	LAUNCH(launch_bucket_insert(t->heads, t->mask, t->next, t->hash, t->max_rows, t->counters, t->overflow,
		sel, inum, (u64*)res, (const u64*)hash, (hipStream_t)stream));
}

int vh_vec_check(vh_table* t, int32_t col, const int32_t* sel, int64_t inum, uint8_t* res,
	const uint64_t* bucket, int32_t key_type, const void* key, void* stream)
{
	int rc = check_table(t, col);
	if (rc)
		return rc;
	if (t->wide[col] || !valid_type(key_type) || inum < 0)
		return fail(VH_ERR_INVALID, "vh_vec_check arguments");
	LAUNCH(launch_check((const i64*)t->cols[col], t->max_rows, sel, inum, res, (const u64*)bucket, key_type,
		key, (hipStream_t)stream));
}

int vh_vec_gather(vh_table* t, int32_t col, const int32_t* sel, int64_t inum, void* res, int32_t res_type,
	const uint64_t* bucket, void* stream)
{
	int rc = check_table(t, col);
	if (rc)
		return rc;
	if (t->wide[col] || !valid_type(res_type) || inum < 0)
		return fail(VH_ERR_INVALID, "vh_vec_gather arguments");
	LAUNCH(launch_gather((const i64*)t->cols[col], t->max_rows, sel, inum, res, res_type, (const u64*)bucket,
		(hipStream_t)stream));
}

int vh_vec_scatter(vh_table* t, int32_t col, const int32_t* sel, int64_t inum, const uint64_t* bucket,
	int32_t val_type, const void* val, void* stream)
{
	int rc = check_table(t, col);
	if (rc)
		return rc;
	if (t->wide[col] || !valid_type(val_type) || inum < 0)
		return fail(VH_ERR_INVALID, "vh_vec_scatter arguments");
	LAUNCH(launch_scatter((i64*)t->cols[col], t->max_rows, sel, inum, (const u64*)bucket, val_type, val,
		(hipStream_t)stream));
}

int vh_vec_aggr_sum(vh_table* t, int32_t col, const int32_t* sel, int64_t inum, const uint64_t* bucket,
	int32_t val_type, const void* val, void* stream)
{
	int rc = check_table(t, col);
	if (rc)
		return rc;
	if (!t->wide[col] || !valid_type(val_type) || !val || inum < 0)
		return fail(VH_ERR_INVALID, "vh_vec_aggr_sum arguments (sum columns are 128-bit)");
	LAUNCH(launch_aggr(t->cols[col], t->max_rows, sel, inum, (const u64*)bucket, val_type, val,
		(hipStream_t)stream));
}

int vh_vec_aggr_count(vh_table* t, int32_t col, const int32_t* sel, int64_t inum, const uint64_t* bucket,
	void* stream)
{
	int rc = check_table(t, col);
	if (rc)
		return rc;
	if (!t->wide[col] || inum < 0)
		return fail(VH_ERR_INVALID, "vh_vec_aggr_count arguments (count columns are 128-bit)");
	LAUNCH(launch_aggr(t->cols[col], t->max_rows, sel, inum, (const u64*)bucket, VH_I64, nullptr,
		(hipStream_t)stream));
}

// bucket_build (kernel_generator.py bucket_build -> ITable::build_index):
// head-insert every row written so far into the index.
int vh_vec_bucket_build(vh_table* t, void* stream)
{
	if (!t)
		return fail(VH_ERR_INVALID, "null table");
	hipStream_t s = (hipStream_t)stream;
	HIP_CALL(launch_fill(t->heads, t->mask + 1, 0, s));
	LAUNCH(launch_index_build(t->heads, t->mask, t->next, t->hash, (u64)t->host_rows, s));
}

// write_pos + write (relalg_translator.cpp:828-907, kernel_generator.py:360-382):
// rows [*base, *base + inum) receive val[sel[k]]; col == -1 is the hash column.
int vh_vec_write_pos(vh_table* t, int64_t inum, int64_t* base)
{
	if (!t || !base || inum < 0)
		return fail(VH_ERR_INVALID, "vh_vec_write_pos arguments");
	if ((u64)(t->host_rows + inum) > t->max_rows)
		return fail(VH_ERR_CAPACITY, "table full");
	*base = t->host_rows;
	t->host_rows += inum;
	return VH_OK;
}

int vh_vec_write(vh_table* t, int32_t col, int64_t base, const int32_t* sel, int64_t inum, int32_t val_type,
	const void* val, void* stream)
{
	if (!t || col < -1 || col >= t->ncols || base < 0 || inum < 0 || (u64)(base + inum) > t->max_rows)
		return fail(VH_ERR_INVALID, "vh_vec_write arguments");
	if (col >= 0 && t->wide[col])
		return fail(VH_ERR_INVALID, "vh_vec_write on a 128-bit column");
	if (inum == 0)
		return VH_OK;
	// a dense write of the selected values into the row range
	u64* dst = col < 0 ? t->hash : t->cols[col];
	LAUNCH(launch_write((i64*)dst + base, sel, inum, val_type, val, (hipStream_t)stream));
}

} // extern "C"
