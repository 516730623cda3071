/*
 * voila_hip.h -- C ABI of the MI355X (gfx950) backend for VOILA's data-parallel
 * operator path: selection, hash-table build/probe with bucket chaining,
 * gather/scatter and hash group-by aggregation.
 *
 * Conventions (stricter than the reference, which aborts on ASSERT):
 *   - every entry point returns an int status, VH_OK == 0;
 *   - device buffers are caller owned unless stated; scratch is library owned
 *     and released by the matching *_free / *_reset;
 *   - `stream` is a hipStream_t passed as void* (NULL = default stream);
 *   - one host thread per handle.
 *
 * Each declaration cites the reference interface it replaces (file:line in
 * t1mm3/vldb_voila).
 */
#ifndef VOILA_HIP_H
#define VOILA_HIP_H

#ifndef __HIPCC_RTC__   /* hipRTC provides these itself */
#include <stddef.h>
#include <stdint.h>
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------ */
/* Types                                                                     */
/* ------------------------------------------------------------------------ */

/* Type codes in the order of the reference's TypeCode (runtime.hpp:270-275). */
enum vh_type {
	VH_INVALID = 0,
	VH_U8, VH_I8, VH_U16, VH_I16, VH_U32, VH_I32, VH_U64, VH_I64, VH_I128
};

enum vh_status {
	VH_OK = 0,
	VH_ERR_INVALID = 1,      /* bad argument / descriptor                  */
	VH_ERR_UNSUPPORTED = 2,  /* shape the backend does not implement       */
	VH_ERR_NOMEM = 3,        /* device allocation failed                   */
	VH_ERR_HIP = 4,          /* HIP runtime error (see vh_last_error)      */
	VH_ERR_CAPACITY = 5,     /* a table or buffer limit was exceeded        */
	VH_ERR_STATE = 6,        /* call out of lifecycle order                */
	VH_ERR_NODEVICE = 7      /* no usable GPU                              */
};

/* Comparison primitives vec_{lt,le,gt,ge,eq,ne} (kernel_generator.py:415-454). */
enum vh_cmp { VH_LT = 0, VH_LE, VH_GT, VH_GE, VH_EQ, VH_NE };

/* Arithmetic primitives vec_{add,sub,mul} (kernel_generator.py:439-445). */
enum vh_expr_op { VH_E_COL = 0, VH_E_CONST, VH_E_ADD, VH_E_SUB, VH_E_MUL };

/* aggr_sum / aggr_count (kernel_generator.py:289-312). */
enum vh_aggr_fn { VH_A_SUM = 0, VH_A_COUNT };

/* relalg::HashJoin::Variant (relalg.hpp:137-140). */
enum vh_join_variant { VH_JOIN01 = 0, VH_JOINN };

#define VH_MAX_COLS     8
#define VH_MAX_PREDS    8
#define VH_MAX_JOINS    4
#define VH_MAX_PAYLOAD  4
#define VH_MAX_JOIN_KEYS 3
#define VH_MAX_KEYS     4
#define VH_MAX_AGGRS    8
#define VH_MAX_EXPR     32

/* A base column: runtime::Attribute {data, type, minmax} (db-engine-paradigms
 * include/common/runtime/Database.hpp:28-75, Types.hpp:11-35).  min/max drive
 * the typing of every derived value, as in typing_pass.cpp:309-326. */
typedef struct vh_column {
	int32_t type;          /* enum vh_type of the stored values            */
	int32_t reserved;
	const void* data;      /* device pointer (GPU entry points)            */
	int64_t min;
	int64_t max;
} vh_column;

/* Conjunctive column-vs-constant predicate: `select(cmp(col, const))`
 * (relalg_translator.cpp:343-378, kernel_generator.py:448-454). */
typedef struct vh_pred {
	int32_t cmp;           /* enum vh_cmp                                  */
	int32_t col;           /* column index in the owning table             */
	int64_t value;
} vh_pred;

/* Expression node.  Nodes form a DAG stored children-first:
 *   VH_E_COL   : a = virtual column index
 *   VH_E_CONST : value
 *   VH_E_ADD/SUB/MUL : a, b = child node indices (< this node's index)   */
typedef struct vh_expr {
	int32_t op;
	int32_t a;
	int32_t b;
	int32_t reserved;
	int64_t value;
} vh_expr;

/* A scanned base table with an optional conjunctive filter. */
typedef struct vh_table_desc {
	int64_t rows;
	int32_t ncols;
	int32_t npreds;
	vh_column cols[VH_MAX_COLS];
	vh_pred preds[VH_MAX_PREDS];
} vh_table_desc;

/* HashJoin of the running probe stream with a build table
 * (relalg.hpp:135-165, relalg_translator.cpp:812-1115).  The build side is
 * materialised (write), indexed (bucket_build) and probed with
 * bucket_lookup / check / gather / bucket_next.  The payload columns become
 * new virtual columns of the probe stream, in join order.  A join on several
 * key columns (the translator's key lists, hashed with rehash) names the
 * first in build_key / probe_key and the others in build_keys / probe_keys;
 * the keys' value ranges together must span at most 32 bits, and so must the
 * payload columns' ranges. */
typedef struct vh_join_desc {
	vh_table_desc build;
	int32_t build_key;                  /* build column of the first key      */
	int32_t probe_key;                  /* virtual column of the first key    */
	int32_t variant;                    /* enum vh_join_variant               */
	int32_t npayload;
	int32_t payload[VH_MAX_PAYLOAD];    /* build column indices               */
	int32_t nkeys;                      /* key columns; 0 and 1 both mean one */
	int32_t build_keys[VH_MAX_JOIN_KEYS - 1];   /* further build key columns  */
	int32_t probe_keys[VH_MAX_JOIN_KEYS - 1];   /* further probe key columns  */
} vh_join_desc;

typedef struct vh_aggr_desc {
	int32_t fn;            /* enum vh_aggr_fn                              */
	int32_t expr;          /* root expression node for VH_A_SUM            */
} vh_aggr_desc;

/*
 * One pipeline: scan(probe) -> select -> join* -> HashAggr(keys, aggregates).
 * Virtual columns: [0, probe.ncols) are the probe columns, followed by the
 * payload columns of joins[0], joins[1], ... in order.
 * nkeys == 0 is HashAggr::Global (relalg.hpp:120-123).
 * Output rows: keys, then aggregates, as relalg_translator.cpp:736-747 emits
 * them; a group exists iff its implicit count > 0 (:698-701).
 */
typedef struct vh_pipeline_desc {
	vh_table_desc probe;
	int32_t njoins;
	int32_t nkeys;
	int32_t naggrs;
	int32_t nexpr;
	vh_join_desc joins[VH_MAX_JOINS];
	int32_t keys[VH_MAX_KEYS];
	vh_aggr_desc aggrs[VH_MAX_AGGRS];
	vh_expr expr[VH_MAX_EXPR];
} vh_pipeline_desc;


/* ------------------------------------------------------------------------ */
/* Variant descriptor: the knobs the autotuner searches.  The reference's    */
/* analog is BlendConfig{computation_type, concurrent_fsms, prefetch}        */
/* (blend_space_point.hpp:10-45).  -1 / 0 select the built-in heuristics.   */
/* ------------------------------------------------------------------------ */

enum vh_group_mode {
	VH_GROUP_AUTO = -1,
	VH_GROUP_REG = 0,      /* <= 8 dense groups in per-lane registers         */
	VH_GROUP_LDS = 1,      /* dense groups in a per-block LDS table           */
	VH_GROUP_GLOBAL = 2,   /* dense groups, global atomics                    */
	VH_GROUP_HASH = 3,     /* packed key, global CAS open addressing          */
	VH_GROUP_LDSHASH = 4,  /* per-block LDS hash table over the dense index,
	                          global atomics when it is full                  */
	VH_GROUP_DENSE = 5,    /* VH_GROUP_REG when it fits, else VH_GROUP_LDS    */
	VH_GROUP_PART = 6      /* dense domain too large for LDS: rows' (group,
	                          values) tuples are partitioned by group range
	                          and each partition aggregated in LDS; no
	                          per-row global atomics                         */
};

/* Physical layout of a join table.  Both give the same matches. */
enum vh_join_layout {
	VH_LAYOUT_AUTO = 0,    /* direct when the key range is dense (range <=
	                          4 x build rows + 65536), else hash            */
	VH_LAYOUT_HASH = 1,    /* open addressing, 64-bit slots (key | payload)   */
	VH_LAYOUT_DIRECT = 2   /* array indexed by key - min holding the payload
	                          in 1, 2 or 4 bytes; repeated keys spill to a
	                          hash table; falls back to hash above 2^30 keys */
};

/* Which probe kernel runs a query.  The generated kernel is HIP source
 * written for the pipeline (column types, operator sequence, join layouts
 * and group-by strategy fixed; pointers and constants read from the plan),
 * compiled with hipRTC for the device at vh_query_new and cached per
 * source text; the generic kernel interprets the same plan.  Results are
 * identical. */
enum vh_kernel {
	VH_KERNEL_AUTO = 0,      /* generated, or generic if hipRTC fails       */
	VH_KERNEL_GENERIC = 1,
	VH_KERNEL_JIT = 2,       /* generated; vh_query_new fails without it.
	                            The next chunk's columns are requested while
	                            this chunk's joins resolve; columns first used
	                            after two filtered joins are read for live
	                            rows only                                   */
	VH_KERNEL_JIT_LAZY = 3,  /* generated, no prefetch; columns first used
	                            after the last filtered join (else after the
	                            predicates) are read for live rows only      */
	VH_KERNEL_JIT_PIPE = 4   /* same kernel as VH_KERNEL_JIT                */
};

typedef struct vh_config {
	int32_t group_mode;    /* enum vh_group_mode                               */
	int32_t grid;          /* probe-kernel workgroups, 0 = auto                */
	int32_t table_shift;   /* join table size vs calc_num_buckets: <<n / >>-n  */
	int32_t tile_shift;    /* batched work item = 16384 << n rows (n < 0: >>); */
	                       /* 8192 << n for a partitioned group-by           */
	int32_t join_layout;   /* enum vh_join_layout                              */
	int32_t kernel;        /* enum vh_kernel (single queries; batches run the
	                          generic kernel)                                  */
	int32_t unroll;        /* rows per lane per step of the generated kernel,
	                          1..16; 0 = by row width (8 for <= 32 bytes of
	                          probe columns, else 2; 4 in 1024-thread
	                          workgroups)                                      */
	int32_t merge;         /* dense LDS group-by, how workgroups combine: 0 =
	                          auto, 1 = partials + a merge launch, 2 = each
	                          workgroup adds into the result with atomics     */
	int32_t part_shift;    /* partition fan-out of the join builds: 0 = auto,
	                          -1 = none (global atomic builds), n > 0 =
	                          partitions of 2^n keys (direct tables, clamped
	                          to 32 KiB slices) or 2^n slots (hash tables,
	                          10..14): fan-out = table entries / 2^n.  Auto
	                          partitions direct tables, and hash tables
	                          larger than one XCD's 4 MiB L2, in 2^13-slot
	                          sub-tables                                     */
	int32_t block;         /* threads per workgroup of the generated probe
	                          kernel (waves per CU = block / 64 x workgroups
	                          per CU): 0 = auto, 256, 512 or 1024            */
	int32_t radix;         /* radix-partitioned probe of a last join whose
	                          hash table is built partitioned: the probe
	                          writes its rows' tuples partitioned, then one
	                          workgroup per partition joins them against the
	                          partition's sub-table in LDS.  1 = on when
	                          it applies; 0 (auto) and -1 = off: the
	                          one-pass probe of the partitioned table was
	                          faster on every shape measured                */
	int32_t reserved;
} vh_config;

/* ------------------------------------------------------------------------ */
/* Runtime: device, memory, errors (plumbing for callers without a HIP       */
/* runtime of their own).                                                     */
/* ------------------------------------------------------------------------ */

const char* vh_last_error(void);
int vh_init(int device);
int vh_malloc(void** ptr, size_t bytes);
int vh_free(void* ptr);
int vh_memcpy_h2d(void* dst, const void* src, size_t bytes);
int vh_memcpy_d2h(void* dst, const void* src, size_t bytes);
int vh_synchronize(void);
int vh_stream_create(void** stream);
int vh_stream_destroy(void* stream);
int vh_event_create(void** event);
int vh_event_destroy(void* event);
int vh_event_record(void* event, void* stream);
int vh_event_elapsed_ms(float* ms, void* start, void* stop);
int vh_stream_synchronize(void* stream);

/* ------------------------------------------------------------------------ */
/* Query lifecycle.  Mirrors the generated library's ABI                     */
/*   voila_query_new / voila_query_run / voila_query_reset / voila_query_free */
/* (emitted at codegen.cpp:371-392, bound by dlsym at compiler.cpp:475-491). */
/* The query owns its hash tables and scratch; the caller owns the columns   */
/* (device pointers in the descriptor, borrowed for the query's lifetime).   */
/* ------------------------------------------------------------------------ */

typedef struct vh_query vh_query;

int vh_query_new(vh_query** q, const vh_pipeline_desc* desc, const vh_config* cfg);
/* Enqueue one full run (clear tables, build every join, probe + aggregate)
 * on `stream`; asynchronous, capturable in a hipGraph. */
int vh_query_run(vh_query* q, void* stream);
/* One phase of vh_query_run, in order: 0 = clear tables and flags,
 * 1 = build every join table (pipelines A+B of relalg_translator.cpp:828-928),
 * 2 = fused probe + aggregate kernel (pipeline C, :933-1111, and the
 * group-by), 3 = merge of per-block partial aggregates (a no-op for the
 * group-by strategies that accumulate globally).  Lets a caller time the
 * probe kernel alone; running 0..3 gives vh_query_run's result. */
int vh_query_run_phase(vh_query* q, void* stream, int phase);
/* Synchronise `stream` and render QueryResult text: one row per group,
 * '|'-separated, numbers as voila_cast<T>::to_cstr prints them
 * (runtime_framework.cpp:56-111, runtime.hpp:348-387).  *len = bytes
 * needed (excluding NUL); VH_ERR_CAPACITY if `cap` is too small. */
int vh_query_result(vh_query* q, void* stream, char* buf, size_t cap, size_t* len,
	int64_t* nrows);
int vh_query_reset(vh_query* q);
int vh_query_free(vh_query* q);
/* Mode the lowering chose (enum vh_group_mode) and the probe-kernel grid. */
int vh_query_info(const vh_query* q, int32_t* group_mode, int32_t* grid, int64_t* groups);
/* Probe kernel in use (VH_KERNEL_GENERIC or VH_KERNEL_JIT) and the generated
 * HIP source (empty for the generic kernel), NUL-terminated into buf when
 * cap > *len; *len = bytes excluding the NUL.  The reference likewise
 * generates C++ per query and compiles it with g++ -O3 -shared
 * (codegen.cpp:211-395, compiler.cpp:19-55). */
int vh_query_kernel(const vh_query* q, int32_t* kernel, char* buf, size_t cap, size_t* len);
/* Events (vh_event_create) that later runs stamp with the generated probe
 * kernel's own start and end (the launch records them itself, so the time
 * between them excludes the dispatch gap that events recorded on the stream
 * around it include); NULL, NULL stops it.  VH_ERR_UNSUPPORTED when the
 * query runs the generic kernel.  No counterpart in the reference, which
 * times whole queries (compiler.cpp:576-588). */
int vh_query_probe_events(vh_query* q, void* start, void* stop);
/* Layout of join j's table: entry_bytes 1, 2 or 4 for a direct table
 * (entries = key range), 8 for hash slots (entries = slots). */
int vh_query_join(const vh_query* q, int32_t j, int32_t* entry_bytes, int64_t* entries);
/* The generated kernel source for a pipeline without creating a query (no
 * device needed): lowering only, same buffer convention. */
int vh_generate_source(const vh_pipeline_desc* desc, const vh_config* cfg, char* buf, size_t cap,
	size_t* len);

/* Type the reference's TypingPass gives [min, max] (typing_pass.cpp:9-22). */
int vh_type_from_minmax(double min, double max);

/* ------------------------------------------------------------------------ */
/* Statement level: the vec_* primitive family (kernel_generator.py:178-660) */
/* over whole device columns.  Calling convention of the generated           */
/* primitives (kernel_generator.py:42-43): sel == NULL means the dense range */
/* [0, inum), else inum ascending int32 positions; maps write res[i] at the  */
/* same position i; selections write a compacted list and its length into   */
/* the device int32 *d_onum.  Row handles are row + 1, 0 = empty, as the     */
/* reference's row pointers (0 = NULL).                                      */
/* ------------------------------------------------------------------------ */

/* Device chained hash table: IHashTable (runtime_struct.hpp:453-503) with
 * NSM rows stored as columns of 64-bit words (128-bit for `wide` columns,
 * which hold sums and counts).  Sized once for max_rows; its bucket array
 * is calc_num_buckets(max_rows) (runtime_struct.cpp:790-814). */
typedef struct vh_table vh_table;
int vh_table_new(vh_table** t, int32_t ncols, const int32_t* wide, int64_t max_rows);
int vh_table_free(vh_table* t);
int vh_table_reset(vh_table* t, void* stream);               /* ITable::reset   */
int vh_table_rows(vh_table* t, void* stream, int64_t* rows);  /* synchronises   */
int vh_table_column(vh_table* t, int32_t col, void** dev_ptr);

/* Selections count set flags per tile into library-owned scratch, one buffer
 * per (device, stream) of the calling thread: selections on different
 * streams run concurrently; two on one stream are ordered by it. */
int vh_vec_seltrue(const int32_t* sel, int64_t inum, int32_t* res, const uint8_t* pred,
	int32_t* d_onum, void* stream);                           /* vec_seltrue  */
int vh_vec_selfalse(const int32_t* sel, int64_t inum, int32_t* res, const uint8_t* pred,
	int32_t* d_onum, void* stream);                           /* vec_selfalse */
/* vh_vec_seltrue without any allocation.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Its tile
 * counts live in a static device array of 512 KiB per GPU (held until the
 * process ends), so inum is at most 131,071 x 4,096 = 536,866,816 positions
 * (VH_ERR_CAPACITY above that).  Every call on one GPU shares that array:
 * calls that overlap on one GPU (other streams, other threads) race; calls
 * ordered by one stream do not.  Same results as vh_vec_seltrue. */
int vh_vec_seltrue_noalloc(const int32_t* sel, int64_t inum, int32_t* res, const uint8_t* pred,
	int32_t* d_onum, void* stream);
/* vec_<cmp>: res[i] = col1[i] cmp (col2 ? col2[i] : constant), both typed `type` */
int vh_vec_cmp(int32_t cmp, int32_t type, const int32_t* sel, int64_t inum, uint8_t* res,
	const void* col1, const void* col2, int64_t constant, void* stream);
/* vec_add/sub/mul: both operands cast to res_type first (kernel_generator.py:439-445) */
int vh_vec_arith(int32_t op, int32_t res_type, const int32_t* sel, int64_t inum, void* res,
	int32_t t1, const void* col1, int32_t t2, const void* col2, int64_t constant, void* stream);
int vh_vec_hash(int32_t type, const int32_t* sel, int64_t inum, uint64_t* res, const void* col,
	void* stream);                                           /* vec_hash     */
int vh_vec_rehash(int32_t type, const int32_t* sel, int64_t inum, uint64_t* res,
	const uint64_t* hash, const void* col, void* stream);    /* vec_rehash   */
int vh_vec_bucket_lookup(vh_table* t, const int32_t* sel, int64_t inum, uint64_t* res,
	const uint64_t* hash, void* stream);                     /* kernel_generator.py:525-541 */
int vh_vec_bucket_next(vh_table* t, const int32_t* sel, int64_t inum, uint64_t* res,
	const uint64_t* bucket, void* stream);                   /* :543-555 */
/* bucket_insert (:470-519): a new row per bucket and call; a tuple whose head
 * already points at a row of this call gets 0 and retries. */
int vh_vec_bucket_insert(vh_table* t, const int32_t* sel, int64_t inum, uint64_t* res,
	const uint64_t* hash, void* stream);
int vh_vec_check(vh_table* t, int32_t col, const int32_t* sel, int64_t inum, uint8_t* res,
	const uint64_t* bucket, int32_t key_type, const void* key, void* stream);   /* :219-244 */
int vh_vec_gather(vh_table* t, int32_t col, const int32_t* sel, int64_t inum, void* res,
	int32_t res_type, const uint64_t* bucket, void* stream);                     /* :322-326 */
int vh_vec_scatter(vh_table* t, int32_t col, const int32_t* sel, int64_t inum,
	const uint64_t* bucket, int32_t val_type, const void* val, void* stream);     /* :346-357 */
int vh_vec_aggr_sum(vh_table* t, int32_t col, const int32_t* sel, int64_t inum,
	const uint64_t* bucket, int32_t val_type, const void* val, void* stream);     /* :289-312 */
int vh_vec_aggr_count(vh_table* t, int32_t col, const int32_t* sel, int64_t inum,
	const uint64_t* bucket, void* stream);
/* write_pos + write (relalg_translator.cpp:828-907): append inum rows;
 * col == -1 writes the hash column. */
int vh_vec_write_pos(vh_table* t, int64_t inum, int64_t* base);
int vh_vec_write(vh_table* t, int32_t col, int64_t base, const int32_t* sel, int64_t inum,
	int32_t val_type, const void* val, void* stream);
/* bucket_build (ITable::build_index, runtime_struct.cpp:1283-1324) */
int vh_vec_bucket_build(vh_table* t, void* stream);

/* Varchar columns: one (1 + max_len)-byte slot per row, a length byte then
 * the bytes (the binary column of Char<n> / Varchar<n>, db-engine-paradigms
 * include/common/runtime/Types.hpp:125-130 and :243-250, as
 * varchar(void* data, idx, max_len) reads it, runtime.hpp:192-205);
 * 1 <= max_len <= 255.  Constants and needles are host memory of at most
 * 255 bytes. */
/* vec_hash on varchar: voila_hash<varchar>::hash, MurmurHash64A with seed 0
 * (runtime.hpp:434-481; kernel_generator.py:465-468) */
int vh_vec_hash_varchar(const int32_t* sel, int64_t inum, uint64_t* res, const uint8_t* col,
	int32_t max_len, void* stream);
/* vec_rehash on varchar: rehash(h, v) = hash(v, seed h) (runtime.hpp:482-484) */
int vh_vec_rehash_varchar(const int32_t* sel, int64_t inum, uint64_t* res, const uint64_t* hash,
	const uint8_t* col, int32_t max_len, void* stream);
/* vec_contains: res[i] = col[i].contains(needle), memmem (runtime.hpp:206-208;
 * kernel_generator.py:435-437) */
int vh_vec_contains(const int32_t* sel, int64_t inum, uint8_t* res, const uint8_t* col, int32_t max_len,
	const char* needle, int32_t needle_len, void* stream);
/* vec_eq / vec_ne on varchar (kernel_generator.py:412-433): varchar::operator==
 * (runtime.hpp:154-156) against col2 (max_len2) or, with col2 NULL, the constant */
int vh_vec_cmp_varchar(int32_t cmp, const int32_t* sel, int64_t inum, uint8_t* res, const uint8_t* col1,
	int32_t max_len1, const uint8_t* col2, int32_t max_len2, const char* constant, int32_t constant_len,
	void* stream);

/* ------------------------------------------------------------------------ */
/* Batched pipelines: many independent pipelines in one persistent launch.   */
/* The reference runs pipelines one at a time, each with its own flavor     */
/* (runtime_framework.cpp:172-198, QueryConfig::pipeline_default_blend at   */
/* runtime_framework.hpp:78); cfgs[i] is pipeline i's variant.              */
/* ------------------------------------------------------------------------ */

typedef struct vh_batch vh_batch;
/* grid: workgroups of each persistent launch (0: eight per CU). */
int vh_batch_new(vh_batch** b, int32_t n, const vh_pipeline_desc* descs, const vh_config* cfgs,
	int32_t grid);
/* Enqueue: one multi-buffer clear + one persistent kernel (asynchronous). */
int vh_batch_run(vh_batch* b, void* stream);
/* Same work, captured once into a hipGraph and replayed (stream != NULL). */
int vh_batch_run_graph(vh_batch* b, void* stream);
int vh_batch_result(vh_batch* b, int32_t i, void* stream, char* buf, size_t cap, size_t* len,
	int64_t* nrows);
int vh_batch_info(const vh_batch* b, int32_t* items, int32_t* grid, int64_t* lds_bytes);
/* Source of the generated kernel a batch runs for this pipeline's probe
 * items (vh_generated_batch: work items from a queue, each item's LDS
 * group-by state flushed to the pipeline's result); no device needed, buffer
 * convention of vh_generate_source.  Pipelines whose sources are equal share
 * one kernel in a batch. */
int vh_generate_batch_source(const vh_pipeline_desc* desc, const vh_config* cfg, char* buf, size_t cap,
	size_t* len);
int vh_batch_free(vh_batch* b);

/* ------------------------------------------------------------------------ */
/* Autotuner: exhaustive timing of the variant space.  The reference's      */
/* explorer compiles and times every flavor and keeps the fastest            */
/* (explorer.cpp:112-427, best pick in SQL at print_base_flavors.py:32-40). */
/* ------------------------------------------------------------------------ */

int vh_variant_count(void);                        /* 48 */
int vh_variant_config(int32_t variant, vh_config* out);
/* ms[i]: fastest of `reps` timed runs of cfgs[i] (full query run on
 * `stream`), or -1 when the variant does not apply to the pipeline. */
int vh_tune_time(const vh_pipeline_desc* desc, const vh_config* cfgs, int32_t n, int32_t reps,
	float* ms, void* stream);
/* Tuning cache: the best variant per pipeline signature (column types,
 * predicate / join / key / sum counts, build sides and probe rows and group
 * domain bucketed by powers of two).  The reference keeps its explorer's
 * timings in SQLite and picks the fastest blend per query with SQL
 * (compiler.cpp:100-122, print_base_flavors.py:32-40); here the pick is kept
 * in memory, and callers persist entries with get / put.
 * vh_tune_best: the cached index (cached = 1), or vh_tune_time over cfgs and
 * the fastest, then cached (cached = 0). */
int vh_tune_signature(const vh_pipeline_desc* desc, char* buf, size_t cap, size_t* len);
int vh_tune_best(const vh_pipeline_desc* desc, const vh_config* cfgs, int32_t n, int32_t reps,
	int32_t* best, int32_t* cached, void* stream);
int vh_tune_cache_get(const char* signature, int32_t* variant);   /* -1 = absent */
int vh_tune_cache_put(const char* signature, int32_t variant);
int vh_tune_cache_clear(void);

/* ------------------------------------------------------------------------ */
/* Learned selector: FP32 MLP 6 -> 64 -> 64 -> 48 trained on the GPU.        */
/* No counterpart in the reference; it replaces the exhaustive search as     */
/* the first guess and falls back to it when its confidence is low.          */
/* ------------------------------------------------------------------------ */

typedef struct vh_mlp vh_mlp;
int vh_mlp_new(vh_mlp** m, int32_t batch_capacity, const float* init_params /* 7,728 floats */);
int vh_mlp_free(vh_mlp* m);
int vh_mlp_params(vh_mlp* m, float* host_params);
/* x [B][6], mask [B][48] (1 = variant applies, NULL = all) and prob_out
 * [B][48] are device buffers. */
int vh_mlp_predict(vh_mlp* m, const float* x, const uint8_t* mask, int32_t B, float* prob_out,
	void* stream);
/* Forward + softmax cross-entropy on labels y [B] + backward; Adam update
 * when lr > 0.  grads_out (host, 7,728 floats) and loss may be NULL. */
int vh_mlp_train_step(vh_mlp* m, const float* x, const uint8_t* mask, const int32_t* y, int32_t B,
	float lr, float* grads_out, float* loss, void* stream);

/* ------------------------------------------------------------------------ */
/* Relational front end ("Flavor::Hip").  Plans are the reference's          */
/* relalg operators (relalg.hpp:12-181) in S-expression form:               */
/*   (hashaggr hash|global (keys k...) (aggrs (sum E) (count) ...) PLAN)    */
/*   (hashjoin join01|joinn LEFT (keys k...) (payl c...) RIGHT (keys k...) (payl c...)) */
/*   (select PRED PLAN)  (project (assign name E)... PLAN)  (scan table c...) */
/* E := column | integer | (+|-|* E E);  PRED := (<|<=|>|>=|=|!= col int)    */
/* | (and PRED PRED).  Columns are named as scanned ("table.col" or "col"). */
/* The catalog mirrors runtime::Database (Database.hpp:28-200): per column  */
/* a type, device data and min/max (Import.cpp:175-186).                    */
/* ------------------------------------------------------------------------ */

typedef struct vh_db vh_db;
int vh_db_new(vh_db** db);
int vh_db_free(vh_db* db);
int vh_db_add_column(vh_db* db, const char* table, const char* column, int32_t type,
	const void* device_data, int64_t rows, int64_t min, int64_t max);
/* Lower a plan to a pipeline descriptor (VH_ERR_UNSUPPORTED for shapes the
 * fused path does not cover). */
int vh_plan_lower(vh_db* db, const char* plan, vh_pipeline_desc* out);
/* vh_query_new on the lowered plan (the caller then runs it as any query). */
int vh_plan_query_new(vh_query** q, vh_db* db, const char* plan, const vh_config* cfg);
/* The VOILA statement sequences the plan becomes and how they are fused. */
int vh_plan_explain(vh_db* db, const char* plan, char* buf, size_t cap, size_t* len);

/* VOILA programs (Program / Pipeline / Lolepop / Statement / Expression /
 * DataStructure, voila.hpp:142-290) in text form:
 *   (program DATA... PIPELINE...)
 *   DATA  (basetable NAME SOURCE (col NAME SOURCE value|key|hash)...)
 *         (table NAME hashtable|table (FLAG...) (col ...)...), FLAG one of
 *         threadlocal readafterwrite flushtomaster
 *   PIPELINE (pipeline (lolepop NAME STMT...)...)
 *   STMT  (assign VAR EXPR PRED) (effect EXPR) (emit EXPR PRED) (loop PRED STMT...)
 *         (wrap PRED STMT...) (blend PRED STMT...) (done) (meta KIND [VAR])
 *   EXPR  (f NAME (ARG...) PRED) (ref NAME) (const TEXT) lolearg lolepred _ (none)
 * The lowering recognises the statement sequences of the relalg translator
 * (scan, select, project, HashJoin materialize / build / probe, HashAggr
 * with flush and re-aggregation; relalg_translator.cpp:202-1115) and maps
 * the query's probe pipeline and its joins' build pipelines onto one
 * pipeline descriptor (VH_ERR_UNSUPPORTED for other shapes).  Base tables
 * resolve against the catalog by their source name.  This is the
 * Flavor::Hip backend's entry point for a typed Program
 * (Compiler::_generate, compiler.cpp:343-385). */
int vh_program_lower(vh_db* db, const char* program, vh_pipeline_desc* out);
int vh_program_query_new(vh_query** q, vh_db* db, const char* program, const vh_config* cfg);

#ifdef __cplusplus
}
#endif

#endif /* VOILA_HIP_H */
