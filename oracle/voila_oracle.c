/*
 * voila_oracle.c -- CPU oracle for VOILA's data-parallel operator path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product loads this library: only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg do, as the
 * checker or as the timed CPU baseline ("port").
 *
 * It holds two independent restatements of the same pipelines
 * (scan -> select -> hash join* -> hash group-by, vh_pipeline_desc):
 *
 *  1. vo_run_pipeline: the reference's vector(1024) flavor, statement by
 *     statement, on 1024-tuple vectors with i32 selection vectors, 16 Ki-tuple
 *     morsels pulled from an atomic cursor, chained NSM hash tables with row
 *     pointers (0 = empty), the bucket_insert conflict rule, thread-local
 *     aggregation flushed to partitions by hash % threads and re-aggregated
 *     (count -> sum).  Each function cites the reference code it restates.
 *  2. vo_bruteforce_pipeline: closed form.  Sorted build keys + binary search,
 *     exact __int128 arithmetic, sort-and-reduce group-by.  Shares no code
 *     with (1) except value loading and text formatting.
 *
 * Both return the reference's QueryResult text: one row per group, columns
 * joined by '|', numbers printed by voila_cast<T>::to_cstr
 * (runtime.hpp:348-387).  Rows are compared sorted (runtime_framework.cpp:238-247).
 *
 * Parity pins (checked in tests/test_oracle.py):
 *   hash/rehash     against the reference's own runtime.hpp (oracle/_ref) and
 *                   exact Python integers;
 *   seltrue/selfalse/selunion against runtime.cpp (oracle/_ref) on random,
 *                   empty, full and sparse selections;
 *   to_cstr         against runtime.hpp (oracle/_ref) and Python str();
 *   varchar hash    MurmurHash64A against SMHasher's published verification
 *                   value and the reference's voila_hash<varchar> (oracle/_ref);
 *   contains / eq   against Python's bytes.find / byte comparison and the
 *                   reference's varchar::contains / operator== (oracle/_ref);
 *   calc_num_buckets against its closed form on small counts;
 *   vo_run_pipeline against vo_bruteforce_pipeline (different algorithm) and
 *                   hand-computed known answers on tiny inputs.
 */
#define _GNU_SOURCE
#include <pthread.h>
#include <stdatomic.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../include/voila_hip.h"

typedef int32_t sel_t;          /* runtime.hpp:235 */
typedef __int128 i128;
typedef unsigned __int128 u128;

/* QueryConfig defaults (runtime_framework.hpp:41-42, :61, :67). */
#define VO_MAX_CARD ((double)(1ull << 40))
#define VO_MIN_BUCKET_COUNT 2048
#define VO_BLOCK_CAPACITY (2u * 1024u * 1024u)   /* ITable::m_block_capacity */

/* ======================================================================== */
/* Types and typing (typing_pass.cpp)                                        */
/* ======================================================================== */

/* numeric_limits<T>::min()/max() as double, TYPE_EXPAND_VOILA_CAST order
 * (runtime.hpp:310-319). */
static const int k_type_order[8] = {VH_U8, VH_I8, VH_U16, VH_I16, VH_U32, VH_I32, VH_U64, VH_I64};

static void ty_limits(int t, double* lo, double* hi)
{
	switch (t) {
	case VH_U8: *lo = 0; *hi = 255.0; break;
	case VH_I8: *lo = -128.0; *hi = 127.0; break;
	case VH_U16: *lo = 0; *hi = 65535.0; break;
	case VH_I16: *lo = -32768.0; *hi = 32767.0; break;
	case VH_U32: *lo = 0; *hi = 4294967295.0; break;
	case VH_I32: *lo = -2147483648.0; *hi = 2147483647.0; break;
	case VH_U64: *lo = 0; *hi = (double)UINT64_MAX; break;
	case VH_I64: *lo = (double)INT64_MIN; *hi = (double)INT64_MAX; break;
	default: *lo = -1e300; *hi = 1e300; break;
	}
}

/* type_from_minmax (typing_pass.cpp:9-22): first of u8,i8,u16,i16,u32,i32,
 * u64,i64 whose limits cover [min,max], else i128. */
int vo_type_from_minmax(double min, double max)
{
	for (int i = 0; i < 8; i++) {
		double lo, hi;
		ty_limits(k_type_order[i], &lo, &hi);
		if (min >= lo && max <= hi)
			return k_type_order[i];
	}
	return VH_I128;
}

/* type_signed_to_unsigned (typing_pass.cpp:89-96) through TYPE_EXPAND_VOILA_CAST:
 * i8->u8, i16->u16, i32->u32; i64 maps to itself (its UTYPE is i64). */
static int ty_scan(int t, int64_t min)
{
	if (min < 0)
		return t;
	switch (t) {
	case VH_I8: return VH_U8;
	case VH_I16: return VH_U16;
	case VH_I32: return VH_U32;
	default: return t;
	}
}

/* Value of a <=64-bit typed canonical int64 widened to i128. */
static inline i128 widen(int t, int64_t v)
{
	return t == VH_U64 ? (i128)(uint64_t)v : (i128)v;
}

/* Reduce an exact value to type t (C conversion semantics: modulo 2^bits). */
static inline i128 wrap_to(int t, i128 v)
{
	switch (t) {
	case VH_U8: return (uint8_t)v;
	case VH_I8: return (int8_t)v;
	case VH_U16: return (uint16_t)v;
	case VH_I16: return (int16_t)v;
	case VH_U32: return (uint32_t)v;
	case VH_I32: return (int32_t)v;
	case VH_U64: return (uint64_t)v;
	case VH_I64: return (int64_t)v;
	default: return v;
	}
}

static inline int64_t load_col(const vh_column* c, int64_t i)
{
	switch (c->type) {
	case VH_U8: return ((const uint8_t*)c->data)[i];
	case VH_I8: return ((const int8_t*)c->data)[i];
	case VH_U16: return ((const uint16_t*)c->data)[i];
	case VH_I16: return ((const int16_t*)c->data)[i];
	case VH_U32: return ((const uint32_t*)c->data)[i];
	case VH_I32: return ((const int32_t*)c->data)[i];
	case VH_U64: return (int64_t)((const uint64_t*)c->data)[i];
	case VH_I64: return ((const int64_t*)c->data)[i];
	default: return 0;
	}
}

typedef struct vo_tp { int type; double dmin, dmax; } vo_tp;   /* ScalarTypeProps */

typedef struct vo_typing {
	int nv;                         /* number of virtual columns          */
	vo_tp vcol[VH_MAX_COLS + VH_MAX_JOINS * VH_MAX_PAYLOAD];
	const vh_column* vsrc[VH_MAX_COLS + VH_MAX_JOINS * VH_MAX_PAYLOAD];
	int vjoin[VH_MAX_COLS + VH_MAX_JOINS * VH_MAX_PAYLOAD];   /* -1 = probe */
	int vpay[VH_MAX_COLS + VH_MAX_JOINS * VH_MAX_PAYLOAD];
	vo_tp node[VH_MAX_EXPR];
	vo_tp aggr[VH_MAX_AGGRS];       /* first-level aggregate column types */
	vo_tp final_aggr[VH_MAX_AGGRS]; /* after the re-aggregation           */
} vo_typing;

static vo_tp tp_from(double lo, double hi)
{
	vo_tp t = { vo_type_from_minmax(lo, hi), lo, hi };
	return t;
}

/* Types every value of the pipeline as TypingPass does: scans from column
 * min/max (:309-326), constants from their value (:465-475), add/sub/mul by
 * the ARITH2_INFIX range rule (:164-176, note: lo/hi products only), sums as
 * max_card * [min,max] of the argument and counts as [0, max_card]
 * (:237-266).  Returns 0 or VH_ERR_INVALID. */
static int vo_type_pipeline(const vh_pipeline_desc* d, vo_typing* ty)
{
	memset(ty, 0, sizeof(*ty));
	int nv = 0;
	for (int c = 0; c < d->probe.ncols; c++) {
		const vh_column* col = &d->probe.cols[c];
		ty->vcol[nv].type = ty_scan(col->type, col->min);
		ty->vcol[nv].dmin = (double)col->min;
		ty->vcol[nv].dmax = (double)col->max;
		ty->vsrc[nv] = col;
		ty->vjoin[nv] = -1;
		ty->vpay[nv] = c;
		nv++;
	}
	for (int j = 0; j < d->njoins; j++) {
		const vh_join_desc* jd = &d->joins[j];
		if (jd->probe_key < 0 || jd->probe_key >= nv)
			return VH_ERR_INVALID;
		for (int p = 0; p < jd->npayload; p++) {
			const vh_column* col = &jd->build.cols[jd->payload[p]];
			ty->vcol[nv].type = ty_scan(col->type, col->min);
			ty->vcol[nv].dmin = (double)col->min;
			ty->vcol[nv].dmax = (double)col->max;
			ty->vsrc[nv] = col;
			ty->vjoin[nv] = j;
			ty->vpay[nv] = p;
			nv++;
		}
	}
	ty->nv = nv;
	for (int k = 0; k < d->nkeys; k++)
		if (d->keys[k] < 0 || d->keys[k] >= nv)
			return VH_ERR_INVALID;

	for (int n = 0; n < d->nexpr; n++) {
		const vh_expr* e = &d->expr[n];
		switch (e->op) {
		case VH_E_COL:
			if (e->a < 0 || e->a >= nv)
				return VH_ERR_INVALID;
			ty->node[n] = ty->vcol[e->a];
			break;
		case VH_E_CONST:
			ty->node[n] = tp_from((double)e->value, (double)e->value);
			break;
		case VH_E_ADD: case VH_E_SUB: case VH_E_MUL: {
			if (e->a < 0 || e->a >= n || e->b < 0 || e->b >= n)
				return VH_ERR_INVALID;
			double alo = ty->node[e->a].dmin, ahi = ty->node[e->a].dmax;
			double blo = ty->node[e->b].dmin, bhi = ty->node[e->b].dmax;
			double x, y;
			if (e->op == VH_E_ADD) { x = alo + blo; y = ahi + bhi; }
			else if (e->op == VH_E_SUB) { x = alo - bhi; y = ahi - blo; }
			else { x = alo * blo; y = ahi * bhi; }
			ty->node[n] = tp_from(x < y ? x : y, x < y ? y : x);
			break;
		}
		default:
			return VH_ERR_INVALID;
		}
	}
	for (int a = 0; a < d->naggrs; a++) {
		const vh_aggr_desc* ad = &d->aggrs[a];
		if (ad->fn == VH_A_SUM) {
			if (ad->expr < 0 || ad->expr >= d->nexpr)
				return VH_ERR_INVALID;
			vo_tp t = ty->node[ad->expr];
			ty->aggr[a] = tp_from(VO_MAX_CARD * t.dmin, VO_MAX_CARD * t.dmax);
		} else if (ad->fn == VH_A_COUNT) {
			ty->aggr[a] = tp_from(0.0, VO_MAX_CARD);
		} else {
			return VH_ERR_INVALID;
		}
		/* re-aggregation: count becomes sum (relalg_translator.cpp:787-800) */
		ty->final_aggr[a] = tp_from(VO_MAX_CARD * ty->aggr[a].dmin,
			VO_MAX_CARD * ty->aggr[a].dmax);
	}
	return 0;
}

int vo_typing_of(const vh_pipeline_desc* d, int32_t* node_types, int32_t* aggr_types,
	int32_t* final_aggr_types)
{
	vo_typing ty;
	int rc = vo_type_pipeline(d, &ty);
	if (rc)
		return rc;
	for (int n = 0; n < d->nexpr; n++)
		node_types[n] = ty.node[n].type;
	for (int a = 0; a < d->naggrs; a++) {
		aggr_types[a] = ty.aggr[a].type;
		final_aggr_types[a] = ty.final_aggr[a].type;
	}
	return 0;
}

/* ======================================================================== */
/* Scalar semantics                                                          */
/* ======================================================================== */

/* voila_hash<T>::hash / rehash (runtime.hpp:413-428): the value is converted
 * to u64 first, sign-extending signed types. */
static inline uint64_t vo_hash_u64(uint64_t x)
{
	x ^= x >> 16;
	x *= UINT64_C(0x45d9f3b);
	x ^= x >> 16;
	x *= UINT64_C(0x45d9f3b);
	x ^= x >> 16;
	return x;
}

uint64_t vo_hash(int64_t canonical_value)
{
	return vo_hash_u64((uint64_t)canonical_value);
}

uint64_t vo_rehash(uint64_t h, int64_t canonical_value)
{
	return vo_hash_u64((uint64_t)canonical_value) ^ (h * UINT64_C(2654435761));
}

/* ======================================================================== */
/* Varchar columns (runtime.hpp:150-224)                                     */
/* ======================================================================== */

/* Slot idx of a varchar column as varchar(void* data, idx, max_len) reads it
 * (runtime.hpp:192-205): (1 + max_len)-byte slots, a u8 length first.  The
 * length is bounded by the slot, as the library bounds it (a well-formed
 * column never exceeds it). */
static const uint8_t* vo_varchar_at(const uint8_t* col, int max_len, int64_t idx, uint32_t* len)
{
	const uint8_t* p = col + (size_t)idx * (size_t)(max_len + 1);
	*len = p[0] < (uint32_t)max_len ? p[0] : (uint32_t)max_len;
	return p + 1;
}

/* voila_hash<varchar>::hash (runtime.hpp:434-481): MurmurHash64A. */
uint64_t vo_murmur64a(const uint8_t* key, size_t len, uint64_t seed)
{
	const uint64_t m = UINT64_C(0xc6a4a7935bd1e995);
	uint64_t h = seed ^ ((uint64_t)len * m);
	size_t i = 0;
	for (; i + 8 <= len; i += 8) {
		uint64_t k;
		memcpy(&k, key + i, 8);          /* little-endian host */
		k *= m;
		k ^= k >> 47;
		k *= m;
		h ^= k;
		h *= m;
	}
	if (len - i) {
		uint64_t t = 0;
		memcpy(&t, key + i, len - i);
		h ^= t;
		h *= m;
	}
	h ^= h >> 47;
	h *= m;
	h ^= h >> 47;
	return h;
}

/* vec_hash / vec_rehash on varchar: hash(v) = MurmurHash64A(v, 0),
 * rehash(h, v) = MurmurHash64A(v, h) (runtime.hpp:482-484). */
void vo_vec_hash_varchar(const sel_t* sel, int64_t inum, uint64_t* res, const uint64_t* h,
	const uint8_t* col, int max_len)
{
	for (int64_t k = 0; k < inum; k++) {
		const int64_t i = sel ? sel[k] : k;
		uint32_t len;
		const uint8_t* s = vo_varchar_at(col, max_len, i, &len);
		res[i] = vo_murmur64a(s, len, h ? h[i] : 0);
	}
}

/* vec_contains (kernel_generator.py:435-437): varchar::contains, memmem
 * (runtime.hpp:206-208). */
void vo_vec_contains(const sel_t* sel, int64_t inum, uint8_t* res, const uint8_t* col, int max_len,
	const uint8_t* needle, int nlen)
{
	for (int64_t k = 0; k < inum; k++) {
		const int64_t i = sel ? sel[k] : k;
		uint32_t len;
		const uint8_t* s = vo_varchar_at(col, max_len, i, &len);
		res[i] = memmem(s, len, needle, (size_t)nlen) != NULL;
	}
}

/* vec_eq / vec_ne on varchar (kernel_generator.py:412-433): varchar::operator==
 * = equal lengths and strncmp(..., len) == 0 (runtime.hpp:154-156). */
void vo_vec_cmp_varchar(int cmp, const sel_t* sel, int64_t inum, uint8_t* res, const uint8_t* col1,
	int max_len1, const uint8_t* col2, int max_len2, const uint8_t* cst, int clen)
{
	for (int64_t k = 0; k < inum; k++) {
		const int64_t i = sel ? sel[k] : k;
		uint32_t la, lb;
		const uint8_t* a = vo_varchar_at(col1, max_len1, i, &la);
		const uint8_t* b = cst;
		lb = (uint32_t)clen;
		if (col2)
			b = vo_varchar_at(col2, max_len2, i, &lb);
		const int eq = la == lb && strncmp((const char*)a, (const char*)b, la) == 0;
		res[i] = (uint8_t)(cmp == VH_EQ ? eq : !eq);
	}
}

/* voila_cast<T>::to_cstr (runtime.hpp:348-387): magnitude digits, then '-'. */
size_t vo_to_cstr_i128(uint64_t lo, uint64_t hi, char* out)
{
	i128 value = (i128)(((u128)hi << 64) | lo);
	char buffer[41];
	u128 tmp = value < 0 ? (u128)(-value) : (u128)value;
	char* d = buffer + sizeof(buffer) - 1;
	*d = 0;
	do {
		--d;
		*d = "0123456789"[(int)(tmp % 10)];
		tmp /= 10;
	} while (tmp != 0);
	if (value < 0) {
		d--;
		*d = '-';
	}
	size_t n = (size_t)((buffer + sizeof(buffer)) - d - 1);
	memcpy(out, d, n);
	out[n] = 0;
	return n;
}

/* calc_num_buckets (runtime_struct.cpp:790-814). */
size_t vo_calc_num_buckets(size_t count, size_t vector_size, size_t min_bucket_count)
{
	size_t bucket_count = count;
	if (min_bucket_count > bucket_count)
		bucket_count = min_bucket_count;
	if (vector_size * 2 > bucket_count)
		bucket_count = vector_size * 2;
	size_t nr = bucket_count;
	size_t exp = 64 - (size_t)__builtin_clzll(nr);
	if (((size_t)1 << exp) < (double)nr / 0.7)
		exp++;
	return (size_t)1 << exp;
}

/* vec_seltrue / vec_selfalse (kernel_generator.py:617-631): ascending indices
 * of sel (or [0,inum) when sel == NULL) whose predicate is (not) set. */
sel_t vo_seltrue(const sel_t* sel, sel_t inum, sel_t* res, const uint8_t* pred)
{
	sel_t onum = 0;
	if (sel) {
		for (sel_t k = 0; k < inum; k++)
			if (pred[sel[k]])
				res[onum++] = sel[k];
	} else {
		for (sel_t k = 0; k < inum; k++)
			if (pred[k])
				res[onum++] = k;
	}
	return onum;
}

sel_t vo_selfalse(const sel_t* sel, sel_t inum, sel_t* res, const uint8_t* pred)
{
	sel_t onum = 0;
	if (sel) {
		for (sel_t k = 0; k < inum; k++)
			if (!pred[sel[k]])
				res[onum++] = sel[k];
	} else {
		for (sel_t k = 0; k < inum; k++)
			if (!pred[k])
				res[onum++] = k;
	}
	return onum;
}

/* ComplexFuncs::selunion (runtime.cpp:67-169): merge of two ascending lists,
 * duplicates dropped.  `result` may alias neither input. */
sel_t vo_selunion(sel_t* result, const sel_t* l, sel_t ln, const sel_t* r, sel_t rn)
{
	sel_t k = 0, i = 0, j = 0;
	while (i < ln && j < rn) {
		if (l[i] < r[j]) result[k++] = l[i++];
		else if (r[j] < l[i]) result[k++] = r[j++];
		else { result[k++] = l[i]; i++; j++; }
	}
	while (i < ln) {
		if (!k || l[i] > result[k - 1]) result[k++] = l[i];
		i++;
	}
	while (j < rn) {
		if (!k || r[j] > result[k - 1]) result[k++] = r[j];
		j++;
	}
	return k;
}

/* ======================================================================== */
/* Blocked row storage and chained hash tables (runtime_struct.{hpp,cpp})   */
/* ======================================================================== */

typedef struct vo_block {
	struct vo_block* next;
	size_t num, cap, width;
	char* data;
} vo_block;

typedef struct vo_space {       /* BlockedSpace */
	vo_block* head;
	vo_block* tail;
	size_t width;
	size_t rows;
} vo_space;

static void space_init(vo_space* s, size_t width)
{
	memset(s, 0, sizeof(*s));
	s->width = width;
}

static void space_free(vo_space* s)
{
	vo_block* b = s->head;
	while (b) {
		vo_block* n = b->next;
		free(b->data);
		free(b);
		b = n;
	}
	s->head = s->tail = NULL;
	s->rows = 0;
}

/* BlockedSpace::append: a block with room for `num` more rows (zeroed, as
 * Block allocates with calloc, runtime_struct.cpp:1015). */
static vo_block* space_append(vo_space* s, size_t num)
{
	if (s->tail && s->tail->num + num <= s->tail->cap)
		return s->tail;
	vo_block* b = (vo_block*)calloc(1, sizeof(vo_block));
	b->cap = num > VO_BLOCK_CAPACITY ? num : VO_BLOCK_CAPACITY;
	b->width = s->width;
	b->data = (char*)calloc(b->cap, s->width);
	if (!b->data) {
		fprintf(stderr, "oracle: out of memory\n");
		abort();
	}
	if (s->tail) s->tail->next = b; else s->head = b;
	s->tail = b;
	return b;
}

/* Row layout (codegen.cpp:67-116 puts keys first, then values, hash, next):
 *   int64 key[nkeys] | i128 aggr[naggr] (group tables) or int64 pay[np] (join
 *   tables) | u64 hash | u64 next */
typedef struct vo_htab {
	vo_space* write;        /* write partitions (1 if thread local)     */
	int nwrite;
	vo_space* flush;        /* flush partitions (bucket_flush)          */
	int nflush;
	uint64_t* heads;        /* hash index: row pointers, 0 = empty      */
	uint64_t mask;
	size_t capacity;
	size_t counter;         /* hash_index_tuple_counter_seq             */
	size_t width, val_off, hash_off, next_off;
	int nkeys, nvals;
	size_t vector_size;
	pthread_mutex_t mutex;
} vo_htab;

static void htab_init(vo_htab* t, int nkeys, int nvals, size_t val_bytes, int nwrite,
	int nflush, size_t vector_size)
{
	memset(t, 0, sizeof(*t));
	t->nkeys = nkeys;
	t->nvals = nvals;
	t->val_off = (8 * (size_t)nkeys + val_bytes - 1) / val_bytes * val_bytes;  /* i128 alignment */
	t->hash_off = t->val_off + val_bytes * (size_t)nvals;
	t->next_off = t->hash_off + 8;
	t->width = t->next_off + 8;
	t->width = (t->width + 15) & ~(size_t)15;
	t->nwrite = nwrite;
	t->write = (vo_space*)calloc((size_t)nwrite, sizeof(vo_space));
	for (int i = 0; i < nwrite; i++)
		space_init(&t->write[i], t->width);
	t->nflush = nflush;
	if (nflush) {
		t->flush = (vo_space*)calloc((size_t)nflush, sizeof(vo_space));
		for (int i = 0; i < nflush; i++)
			space_init(&t->flush[i], t->width);
	}
	t->vector_size = vector_size;
	pthread_mutex_init(&t->mutex, NULL);
}

static void htab_free(vo_htab* t)
{
	for (int i = 0; i < t->nwrite; i++)
		space_free(&t->write[i]);
	for (int i = 0; i < t->nflush; i++)
		space_free(&t->flush[i]);
	free(t->write);
	free(t->flush);
	free(t->heads);
	pthread_mutex_destroy(&t->mutex);
}

#define ROW_HASH(t, r) (*(uint64_t*)((char*)(r) + (t)->hash_off))
#define ROW_NEXT(t, r) (*(uint64_t*)((char*)(r) + (t)->next_off))
#define ROW_KEY(t, r, k) (((int64_t*)(r))[k])

/* remove_hash_index (runtime_struct.cpp:1265-1280). */
static void htab_remove_index(vo_htab* t)
{
	free(t->heads);
	t->heads = NULL;
	t->mask = 0;
	t->capacity = 0;
	t->counter = 0;
}

/* Sequential head insertion of every row of a space
 * (reinsert_hash_buckets_sequential, runtime_struct.cpp:174-211). */
static void htab_reinsert_space(vo_htab* t, vo_space* s)
{
	for (vo_block* b = s->head; b; b = b->next)
		for (size_t i = 0; i < b->num; i++) {
			char* row = b->data + i * t->width;
			uint64_t idx = ROW_HASH(t, row) & t->mask;
			ROW_NEXT(t, row) = t->heads[idx];
			t->heads[idx] = (uint64_t)(uintptr_t)row;
		}
}

/* Parallel CAS head insertion of one thread's partition
 * (reinsert_hash_buckets_parallel, runtime_struct.cpp:322-356). */
static void htab_reinsert_space_cas(vo_htab* t, vo_space* s)
{
	for (vo_block* b = s->head; b; b = b->next)
		for (size_t i = 0; i < b->num; i++) {
			char* row = b->data + i * t->width;
			uint64_t idx = ROW_HASH(t, row) & t->mask;
			_Atomic uint64_t* head = (_Atomic uint64_t*)&t->heads[idx];
			uint64_t old = atomic_load_explicit(head, memory_order_relaxed);
			do {
				ROW_NEXT(t, row) = old;
			} while (!atomic_compare_exchange_weak(head, &old, (uint64_t)(uintptr_t)row));
		}
}

static size_t htab_rows(const vo_htab* t)
{
	size_t c = 0;
	for (int i = 0; i < t->nwrite; i++)
		for (vo_block* b = t->write[i].head; b; b = b->next)
			c += b->num;
	return c;
}

/* ITable::build_index for a thread-local table (runtime_struct.cpp:1283-1324):
 * size by calc_num_buckets(rows); keep the index unless forced or resized. */
static void htab_build_index_local(vo_htab* t, int force)
{
	size_t count = htab_rows(t);
	size_t nb = vo_calc_num_buckets(count, t->vector_size, VO_MIN_BUCKET_COUNT);
	uint64_t mask = nb - 1;
	if (!force && t->heads && mask == t->mask)
		return;
	htab_remove_index(t);
	t->mask = mask;
	t->capacity = nb;
	t->heads = (uint64_t*)calloc(nb, sizeof(uint64_t));
	for (int i = 0; i < t->nwrite; i++)
		htab_reinsert_space(t, &t->write[i]);
}

/* ======================================================================== */
/* Pipeline execution, vector(1024) flavor                                   */
/* ======================================================================== */

typedef struct vo_query vo_query;

typedef struct vo_level {         /* per-join scratch vectors */
	uint64_t* bucket;
	uint64_t* hash;
	sel_t* active;
	sel_t* active2;
	sel_t* hit;
	uint8_t* match;
} vo_level;

typedef struct vo_thread {
	vo_query* q;
	int tid;
	vo_htab gt;                  /* thread-local group table (first level)  */
	vo_htab rt;                  /* thread-local re-aggregation table       */
	int64_t* v[VH_MAX_COLS + VH_MAX_JOINS * VH_MAX_PAYLOAD];  /* vcol vectors */
	i128* node[VH_MAX_EXPR];
	vo_level lvl[VH_MAX_JOINS];
	sel_t* sel_a;
	sel_t* sel_b;
	uint8_t* pred;
	/* group-by scratch */
	uint64_t* g_hash;
	uint64_t* g_bucket;
	uint64_t* g_newpos;
	sel_t* g_miss;
	sel_t* g_miss2;
	sel_t* g_hit;
	sel_t* g_hit2;
	sel_t* g_found;
	sel_t* g_tmp;
	uint8_t* g_pred;
	uint8_t* g_eq;
	int64_t* keyv[VH_MAX_KEYS];
	i128* aggv[VH_MAX_AGGRS];
	char* out;
	size_t out_len, out_cap;
	size_t out_rows;
} vo_thread;

struct vo_query {
	const vh_pipeline_desc* d;
	vo_typing ty;
	int nthreads;
	size_t vsize;
	size_t morsel;
	vo_htab jt[VH_MAX_JOINS];
	_Atomic int64_t cursor;
	pthread_barrier_t barrier;
	vo_thread* th;
};

static int probe_pass(const vh_table_desc* tb, int64_t row)
{
	for (int p = 0; p < tb->npreds; p++) {
		const vh_pred* pr = &tb->preds[p];
		const vh_column* c = &tb->cols[pr->col];
		i128 a = widen(ty_scan(c->type, c->min), load_col(c, row));
		i128 b = pr->value;
		int ok;
		switch (pr->cmp) {
		case VH_LT: ok = a < b; break;
		case VH_LE: ok = a <= b; break;
		case VH_GT: ok = a > b; break;
		case VH_GE: ok = a >= b; break;
		case VH_EQ: ok = a == b; break;
		default: ok = a != b; break;
		}
		if (!ok)
			return 0;
	}
	return 1;
}

/* select(cmp(col, const)) per conjunct: vec_<cmp> into a u8 predicate vector
 * (kernel_generator.py:448-454) then vec_seltrue (:617-631). */
static sel_t vo_select_vector(const vh_table_desc* tb, int64_t base, sel_t n,
	sel_t* sel_out, sel_t* tmp, uint8_t* pred)
{
	sel_t* cur = NULL;
	sel_t num = n;
	for (int p = 0; p < tb->npreds; p++) {
		const vh_pred* pr = &tb->preds[p];
		const vh_column* c = &tb->cols[pr->col];
		int t = ty_scan(c->type, c->min);
		i128 b = pr->value;
		for (sel_t k = 0; k < num; k++) {
			sel_t i = cur ? cur[k] : k;
			i128 a = widen(t, load_col(c, base + i));
			int ok;
			switch (pr->cmp) {
			case VH_LT: ok = a < b; break;
			case VH_LE: ok = a <= b; break;
			case VH_GT: ok = a > b; break;
			case VH_GE: ok = a >= b; break;
			case VH_EQ: ok = a == b; break;
			default: ok = a != b; break;
			}
			pred[i] = (uint8_t)ok;
		}
		sel_t* dst = (cur == sel_out) ? tmp : sel_out;
		num = vo_seltrue(cur, num, dst, pred);
		cur = dst;
	}
	if (cur && cur != sel_out) {
		memcpy(sel_out, cur, (size_t)num * sizeof(sel_t));
	}
	if (!cur) {
		for (sel_t k = 0; k < num; k++)
			sel_out[k] = k;
	}
	return num;
}

/* ---- group-by on one batch (relalg_translator.cpp:582-656) -------------- */

static int rows_keys_equal(vo_htab* t, const char* row, int64_t* const* keyv, sel_t i)
{
	for (int k = 0; k < t->nkeys; k++)
		if (ROW_KEY(t, row, k) != keyv[k][i])
			return 0;
	return 1;
}

/* vec_bucket_insert (kernel_generator.py:470-519) with hash_append_prealloc /
 * prune (runtime_struct.hpp:458-485) and get_current_block_range
 * (runtime_struct.cpp:1416-1435).  A tuple whose bucket head already points
 * into this call's new-row range is a conflict: res = 0, it retries. */
static void vo_bucket_insert(vo_htab* t, const sel_t* sel, sel_t inum,
	uint64_t* res, const uint64_t* hash)
{
	if (!inum)
		return;
	if ((t->counter + (size_t)inum) * 2 > t->capacity)
		htab_build_index_local(t, 0);
	vo_block* block = space_append(&t->write[0], (size_t)inum);
	uint64_t ptr_begin = (uint64_t)(uintptr_t)(block->data + block->num * t->width);
	uint64_t ptr_end = ptr_begin + (uint64_t)inum * t->width;
	size_t ins = 0;
	for (sel_t k = 0; k < inum; k++) {
		sel_t i = sel[k];
		uint64_t idx = hash[i] & t->mask;
		uint64_t head = t->heads[idx];
		res[i] = 0;
		int conflict = head >= ptr_begin && head <= ptr_end;
		if (!conflict) {
			char* row = (char*)(uintptr_t)(ptr_begin + t->width * ins);
			res[i] = (uint64_t)(uintptr_t)row;
			ROW_NEXT(t, row) = head;
			t->heads[idx] = res[i];
			ins++;
		}
	}
	t->counter += ins;
	block->num += ins;
}

static void vo_groupby_batch(vo_thread* th, vo_htab* t, const sel_t* sel, sel_t num,
	int64_t* const* keyv, i128* const* aggv, const uint8_t* is_count)
{
	if (!num)
		return;
	uint64_t* hash = th->g_hash;
	uint64_t* bucket = th->g_bucket;
	uint64_t* new_pos = th->g_newpos;
	sel_t* miss = th->g_miss;
	sel_t* miss2 = th->g_miss2;
	sel_t* hit = th->g_hit;
	sel_t* hit2 = th->g_hit2;
	sel_t* found = th->g_found;
	sel_t* tmp = th->g_tmp;
	uint8_t* eq = th->g_eq;
	uint8_t* equal = th->g_pred;

	/* hash = hash(k0) [rehash(hash, k1) ...] | sel  (vec_hash / vec_rehash) */
	for (sel_t k = 0; k < num; k++) {
		sel_t i = sel[k];
		uint64_t h = 42;               /* kGlobalAggrHashVal, runtime.hpp:741 */
		if (t->nkeys) {
			h = vo_hash(keyv[0][i]);
			for (int c = 1; c < t->nkeys; c++)
				h = vo_rehash(h, keyv[c][i]);
		}
		hash[i] = h;
	}
	memcpy(miss, sel, (size_t)num * sizeof(sel_t));
	sel_t miss_num = num;

	while (miss_num > 0) {
		/* bucket = bucket_lookup(ht, hash) | miss; empty = eq(bucket, 0) */
		for (sel_t k = 0; k < miss_num; k++) {
			sel_t i = miss[k];
			bucket[i] = t->heads[hash[i] & t->mask];
			eq[i] = bucket[i] == 0;
		}
		sel_t hit_num = vo_selfalse(miss, miss_num, hit, eq);
		miss_num = vo_seltrue(miss, miss_num, miss2, eq);
		{ sel_t* x = miss; miss = miss2; miss2 = x; }

		while (hit_num > 0) {
			/* equal = check(key0) and check(key1) ... | hit */
			for (sel_t k = 0; k < hit_num; k++) {
				sel_t i = hit[k];
				equal[i] = (uint8_t)rows_keys_equal(t, (const char*)(uintptr_t)bucket[i], keyv, i);
			}
			sel_t found_num = vo_seltrue(hit, hit_num, found, equal);
			/* aggr_sum / aggr_count | found */
			for (sel_t k = 0; k < found_num; k++) {
				sel_t i = found[k];
				i128* acc = (i128*)((char*)(uintptr_t)bucket[i] + t->val_off);
				for (int a = 0; a < t->nvals; a++)
					acc[a] += is_count[a] ? 1 : aggv[a][i];
			}
			/* hit = selfalse(equal) | hit; bucket = bucket_next | hit */
			sel_t h2 = vo_selfalse(hit, hit_num, hit2, equal);
			for (sel_t k = 0; k < h2; k++) {
				sel_t i = hit2[k];
				bucket[i] = ROW_NEXT(t, (const char*)(uintptr_t)bucket[i]);
				eq[i] = bucket[i] == 0;
			}
			/* miss = selunion(miss, seltrue(empty) | hit); hit = selfalse(empty) | hit */
			sel_t e_num = vo_seltrue(hit2, h2, tmp, eq);
			sel_t u = vo_selunion(miss2, miss, miss_num, tmp, e_num);
			{ sel_t* x = miss; miss = miss2; miss2 = x; }
			miss_num = u;
			hit_num = vo_selfalse(hit2, h2, hit, eq);
		}

		/* new_pos = bucket_insert(ht, hash) | miss */
		vo_bucket_insert(t, miss, miss_num, new_pos, hash);
		/* can_scatter = selfalse(eq(new_pos, 0)) | miss; scatter keys, hash */
		for (sel_t k = 0; k < miss_num; k++) {
			sel_t i = miss[k];
			if (new_pos[i]) {
				char* row = (char*)(uintptr_t)new_pos[i];
				for (int c = 0; c < t->nkeys; c++)
					ROW_KEY(t, row, c) = keyv[c][i];
				ROW_HASH(t, row) = hash[i];
			}
		}
	}
}

/* ---- output (QueryResult, runtime_framework.cpp:56-111) ------------------ */

static void out_push(vo_thread* th, const char* s, size_t n)
{
	if (th->out_len + n + 1 > th->out_cap) {
		size_t cap = th->out_cap ? th->out_cap * 2 : 4096;
		while (cap < th->out_len + n + 1)
			cap *= 2;
		th->out = (char*)realloc(th->out, cap);
		th->out_cap = cap;
	}
	memcpy(th->out + th->out_len, s, n);
	th->out_len += n;
	th->out[th->out_len] = 0;
}

static void out_value(vo_thread* th, int first, i128 v)
{
	char buf[48];
	size_t n = vo_to_cstr_i128((uint64_t)(u128)v, (uint64_t)((u128)v >> 64), buf);
	if (!first)
		out_push(th, "|", 1);
	out_push(th, buf, n);
}

/* ---- expression evaluation: vec_add/sub/mul cast both operands to the
 * result type first (kernel_generator.py:439-445) ------------------------- */

static void eval_nodes(vo_thread* th, const sel_t* sel, sel_t num)
{
	const vh_pipeline_desc* d = th->q->d;
	const vo_typing* ty = &th->q->ty;
	for (int n = 0; n < d->nexpr; n++) {
		const vh_expr* e = &d->expr[n];
		i128* r = th->node[n];
		int rt = ty->node[n].type;
		switch (e->op) {
		case VH_E_COL: {
			int ct = ty->vcol[e->a].type;
			const int64_t* v = th->v[e->a];
			for (sel_t k = 0; k < num; k++) { sel_t i = sel[k]; r[i] = widen(ct, v[i]); }
			break;
		}
		case VH_E_CONST:
			for (sel_t k = 0; k < num; k++) r[sel[k]] = e->value;
			break;
		default: {
			const i128* a = th->node[e->a];
			const i128* b = th->node[e->b];
			for (sel_t k = 0; k < num; k++) {
				sel_t i = sel[k];
				i128 x = wrap_to(rt, a[i]), y = wrap_to(rt, b[i]);
				i128 z = e->op == VH_E_ADD ? x + y : e->op == VH_E_SUB ? x - y : x * y;
				r[i] = wrap_to(rt, z);
			}
			break;
		}
		}
	}
}

/* Group-by stage of the first aggregation: keys from virtual columns, sums of
 * expression roots, counts. */
static void stage_groupby(vo_thread* th, const sel_t* sel, sel_t num)
{
	const vh_pipeline_desc* d = th->q->d;
	if (!num)
		return;
	eval_nodes(th, sel, num);
	int64_t* keyv[VH_MAX_KEYS];
	i128* aggv[VH_MAX_AGGRS + 1];
	uint8_t is_count[VH_MAX_AGGRS + 1];
	for (int k = 0; k < d->nkeys; k++)
		keyv[k] = th->v[d->keys[k]];
	int nv = 0;
	for (int a = 0; a < d->naggrs; a++) {
		is_count[nv] = d->aggrs[a].fn == VH_A_COUNT;
		aggv[nv++] = is_count[a] ? NULL : th->node[d->aggrs[a].expr];
	}
	is_count[nv] = 1;   /* implicit count column (relalg_translator.cpp:469-472) */
	aggv[nv++] = NULL;
	vo_groupby_batch(th, &th->gt, sel, num, keyv, aggv, is_count);
}

static int join_nkeys(const vh_join_desc* jd) { return jd->nkeys > 1 ? jd->nkeys : 1; }
static int join_build_key(const vh_join_desc* jd, int k) { return k ? jd->build_keys[k - 1] : jd->build_key; }
static int join_probe_key(const vh_join_desc* jd, int k) { return k ? jd->probe_keys[k - 1] : jd->probe_key; }

/* Probe stage of join j (relalg_translator.cpp:933-1111, §3.3):
 * bucket_lookup, check, gather, emit, [Join01 stop], bucket_next. */
static void stage_join(vo_thread* th, int j, const sel_t* sel, sel_t num)
{
	vo_query* q = th->q;
	const vh_pipeline_desc* d = q->d;
	if (j == d->njoins) {
		stage_groupby(th, sel, num);
		return;
	}
	if (!num)
		return;
	const vh_join_desc* jd = &d->joins[j];
	vo_htab* t = &q->jt[j];
	vo_level* L = &th->lvl[j];
	const int nk = join_nkeys(jd);
	const int64_t* keyv[VH_MAX_JOIN_KEYS];
	for (int k = 0; k < nk; k++)
		keyv[k] = th->v[join_probe_key(jd, k)];
	int first_pay = q->d->probe.ncols;
	for (int jj = 0; jj < j; jj++)
		first_pay += d->joins[jj].npayload;

	/* bucket = bucket_lookup(ht, hash(key)); active = selfalse(eq(0, bucket)) */
	for (sel_t k = 0; k < num; k++) {
		sel_t i = sel[k];
		uint64_t h = vo_hash(keyv[0][i]);
		for (int c = 1; c < nk; c++)
			h = vo_rehash(h, keyv[c][i]);
		L->hash[i] = h;
		L->bucket[i] = t->heads[L->hash[i] & t->mask];
		L->match[i] = L->bucket[i] == 0;
	}
	sel_t active_num = vo_selfalse(sel, num, L->active, L->match);
	sel_t* active = L->active;
	sel_t* active2 = L->active2;

	while (active_num > 0) {
		/* match = check(ht.key, bucket, key) | active; hit = seltrue(match) */
		for (sel_t k = 0; k < active_num; k++) {
			sel_t i = active[k];
			const char* row = (const char*)(uintptr_t)L->bucket[i];
			int eq = 1;
			for (int c = 0; c < nk; c++)
				eq = eq && ROW_KEY(t, row, c) == keyv[c][i];
			L->match[i] = (uint8_t)eq;
		}
		sel_t hit_num = vo_seltrue(active, active_num, L->hit, L->match);
		if (hit_num) {
			/* payload = gather(ht.col, bucket) | hit; emit */
			for (int p = 0; p < jd->npayload; p++) {
				int64_t* dst = th->v[first_pay + p];
				for (sel_t k = 0; k < hit_num; k++) {
					sel_t i = L->hit[k];
					dst[i] = ((const int64_t*)(uintptr_t)L->bucket[i])[nk + p];
				}
			}
			stage_join(th, j + 1, L->hit, hit_num);
		}
		if (jd->variant == VH_JOIN01) {
			/* active = selfalse(match) | active (:1032-1042) */
			active_num = vo_selfalse(active, active_num, active2, L->match);
			sel_t* x = active; active = active2; active2 = x;
		}
		/* bucket = bucket_next(ht, bucket) | active; active = selfalse(eq(bucket, 0)) */
		for (sel_t k = 0; k < active_num; k++) {
			sel_t i = active[k];
			L->bucket[i] = ROW_NEXT(t, (const char*)(uintptr_t)L->bucket[i]);
			L->match[i] = L->bucket[i] == 0;
		}
		active_num = vo_selfalse(active, active_num, active2, L->match);
		sel_t* x = active; active = active2; active2 = x;
	}
}

static int64_t next_morsel(vo_query* q, int64_t rows, int64_t* n)
{
	int64_t off = atomic_fetch_add(&q->cursor, (int64_t)q->morsel);
	if (off >= rows)
		return -1;
	*n = rows - off < (int64_t)q->morsel ? rows - off : (int64_t)q->morsel;
	return off;
}

/* Build-side pipelines: materialise (write_pos/write, relalg_translator.cpp:
 * 828-907) into the thread's write partition, then bucket_build
 * (:914-928 -> ITable::build_index, runtime_struct.cpp:1283-1324) with CAS
 * head insertion of each partition. */
static void run_build(vo_thread* th, int j)
{
	vo_query* q = th->q;
	const vh_join_desc* jd = &q->d->joins[j];
	const vh_table_desc* tb = &jd->build;
	vo_htab* t = &q->jt[j];
	vo_space* sp = &t->write[th->tid];
	int64_t n;
	int64_t off;
	while ((off = next_morsel(q, tb->rows, &n)) >= 0) {
		for (int64_t v0 = 0; v0 < n; v0 += (int64_t)q->vsize) {
			sel_t vn = (sel_t)((n - v0) < (int64_t)q->vsize ? (n - v0) : (int64_t)q->vsize);
			int64_t base = off + v0;
			sel_t num = vo_select_vector(tb, base, vn, th->sel_a, th->sel_b, th->pred);
			if (!num)
				continue;
			vo_block* b = space_append(sp, (size_t)num);
			char* w = b->data + b->num * t->width;
			for (sel_t k = 0; k < num; k++) {
				sel_t i = th->sel_a[k];
				char* row = w + (size_t)k * t->width;
				/* keys (hash, then rehash per further column), then payloads */
				const int nk = join_nkeys(jd);
				uint64_t h = 0;
				for (int k = 0; k < nk; k++) {
					int64_t key = load_col(&tb->cols[join_build_key(jd, k)], base + i);
					ROW_KEY(t, row, k) = key;
					h = k ? vo_rehash(h, key) : vo_hash(key);
				}
				for (int p = 0; p < jd->npayload; p++)
					((int64_t*)row)[nk + p] = load_col(&tb->cols[jd->payload[p]], base + i);
				ROW_HASH(t, row) = h;
			}
			b->num += (size_t)num;
		}
	}
	pthread_barrier_wait(&q->barrier);
	if (th->tid == 0) {
		size_t count = htab_rows(t);
		size_t nb = vo_calc_num_buckets(count, q->vsize, VO_MIN_BUCKET_COUNT);
		t->mask = nb - 1;
		t->capacity = nb;
		t->heads = (uint64_t*)calloc(nb, sizeof(uint64_t));
		atomic_store(&q->cursor, 0);
	}
	pthread_barrier_wait(&q->barrier);
	htab_reinsert_space_cas(t, sp);
	pthread_barrier_wait(&q->barrier);
}

static void run_probe(vo_thread* th)
{
	vo_query* q = th->q;
	const vh_pipeline_desc* d = q->d;
	int64_t n, off;
	while ((off = next_morsel(q, d->probe.rows, &n)) >= 0) {
		for (int64_t v0 = 0; v0 < n; v0 += (int64_t)q->vsize) {
			sel_t vn = (sel_t)((n - v0) < (int64_t)q->vsize ? (n - v0) : (int64_t)q->vsize);
			int64_t base = off + v0;
			sel_t num = vo_select_vector(&d->probe, base, vn, th->sel_a, th->sel_b, th->pred);
			if (!num)
				continue;
			/* scan: the referenced probe columns of the selected tuples */
			for (int c = 0; c < d->probe.ncols; c++) {
				int64_t* dst = th->v[c];
				const vh_column* col = &d->probe.cols[c];
				for (sel_t k = 0; k < num; k++) {
					sel_t i = th->sel_a[k];
					dst[i] = load_col(col, base + i);
				}
			}
			stage_join(th, 0, th->sel_a, num);
		}
	}
}

/* bucket_flush (ITable::flush2partitions, runtime_struct.cpp:1326-1350 and
 * runtime_partition hash % parts, runtime_utils.cpp:29-44). */
static void flush_to_partitions(vo_htab* t)
{
	for (vo_block* b = t->write[0].head; b; b = b->next)
		for (size_t i = 0; i < b->num; i++) {
			char* row = b->data + i * t->width;
			int p = (int)(ROW_HASH(t, row) % (uint64_t)t->nflush);
			vo_block* db = space_append(&t->flush[p], 1);
			memcpy(db->data + db->num * t->width, row, t->width);
			db->num++;
		}
	space_free(&t->write[0]);
	htab_remove_index(t);
}

/* Re-aggregation of partition `tid` of every thread table: read morsels with
 * count > 0 (relalg_translator.cpp:692-771), HashAggr with count -> sum
 * (:776-808), then read with count > 0 and emit the output rows. */
static void run_reaggregate(vo_thread* th)
{
	vo_query* q = th->q;
	const vh_pipeline_desc* d = q->d;
	int nk = d->nkeys, na = d->naggrs;
	vo_htab* gt0 = &q->th[0].gt;
	uint8_t is_count[VH_MAX_AGGRS + 1];
	i128* aggv[VH_MAX_AGGRS + 1];
	for (int a = 0; a <= na; a++) {
		is_count[a] = a == na;
		aggv[a] = th->aggv[a < VH_MAX_AGGRS ? a : 0];
	}
	aggv[na] = NULL;
	sel_t* sel = th->sel_a;
	for (int src = 0; src < q->nthreads; src++) {
		vo_htab* t = &q->th[src].gt;
		for (vo_block* b = t->flush[th->tid].head; b; b = b->next) {
			for (size_t v0 = 0; v0 < b->num; v0 += q->vsize) {
				size_t vn = b->num - v0 < q->vsize ? b->num - v0 : q->vsize;
				sel_t num = 0;
				for (size_t r = 0; r < vn; r++) {
					const char* row = b->data + (v0 + r) * t->width;
					const i128* acc = (const i128*)(row + t->val_off);
					if (acc[na] > 0) {             /* gt(read(count), 0) */
						sel_t i = (sel_t)r;
						for (int k = 0; k < nk; k++)
							th->keyv[k][i] = ROW_KEY(t, row, k);
						for (int a = 0; a < na; a++)
							th->aggv[a][i] = acc[a];
						sel[num++] = i;
					}
				}
				vo_groupby_batch(th, &th->rt, sel, num, th->keyv, aggv, is_count);
			}
		}
	}
	(void)gt0;
	/* final read: count > 0 -> output rows */
	vo_htab* rt = &th->rt;
	for (vo_block* b = rt->write[0].head; b; b = b->next)
		for (size_t r = 0; r < b->num; r++) {
			const char* row = b->data + r * rt->width;
			const i128* acc = (const i128*)(row + rt->val_off);
			if (!(acc[na] > 0))
				continue;
			for (int k = 0; k < nk; k++)
				out_value(th, k == 0, widen(q->ty.vcol[d->keys[k]].type, ROW_KEY(rt, row, k)));
			for (int a = 0; a < na; a++)
				out_value(th, nk == 0 && a == 0, wrap_to(q->ty.final_aggr[a].type, acc[a]));
			out_push(th, "\n", 1);
			th->out_rows++;
		}
}

static void* worker(void* arg)
{
	vo_thread* th = (vo_thread*)arg;
	vo_query* q = th->q;
	for (int j = 0; j < q->d->njoins; j++) {
		run_build(th, j);
	}
	run_probe(th);
	pthread_barrier_wait(&q->barrier);
	flush_to_partitions(&th->gt);
	pthread_barrier_wait(&q->barrier);
	run_reaggregate(th);
	return NULL;
}

static void* xalloc(size_t n)
{
	void* p = calloc(1, n ? n : 1);
	if (!p) {
		fprintf(stderr, "oracle: out of memory\n");
		abort();
	}
	return p;
}

/*
 * Runs the pipeline as the vector(vector_size) flavor with `nthreads` workers
 * and morsels of `morsel_size` tuples.  Output: QueryResult text into `out`
 * (capacity `cap`); *len = bytes needed.  *elapsed_ns = wall time of the
 * query run (table set-up excluded, as Query::run is timed after
 * voila_query_new, compiler.cpp:521-547).
 */
int vo_run_pipeline(const vh_pipeline_desc* d, int nthreads, int vector_size,
	int morsel_size, char* out, size_t cap, size_t* len, int64_t* nrows,
	int64_t* elapsed_ns)
{
	if (nthreads < 1 || vector_size < 1 || morsel_size < vector_size)
		return VH_ERR_INVALID;
	if (d->nkeys > VH_MAX_KEYS || d->naggrs > VH_MAX_AGGRS || d->njoins > VH_MAX_JOINS ||
		d->nexpr > VH_MAX_EXPR || d->probe.ncols > VH_MAX_COLS)
		return VH_ERR_INVALID;
	vo_query q;
	memset(&q, 0, sizeof(q));
	q.d = d;
	q.nthreads = nthreads;
	q.vsize = (size_t)vector_size;
	q.morsel = (size_t)morsel_size;
	int rc = vo_type_pipeline(d, &q.ty);
	if (rc)
		return rc;
	for (int j = 0; j < d->njoins; j++) {
		if (d->joins[j].npayload > VH_MAX_PAYLOAD)
			return VH_ERR_INVALID;
		htab_init(&q.jt[j], join_nkeys(&d->joins[j]), d->joins[j].npayload, 8, nthreads, 0, q.vsize);
	}
	pthread_barrier_init(&q.barrier, NULL, (unsigned)nthreads);
	q.th = (vo_thread*)xalloc(sizeof(vo_thread) * (size_t)nthreads);
	size_t vs = q.vsize;
	for (int t = 0; t < nthreads; t++) {
		vo_thread* th = &q.th[t];
		th->q = &q;
		th->tid = t;
		htab_init(&th->gt, d->nkeys, d->naggrs + 1, 16, 1, nthreads, vs);
		htab_build_index_local(&th->gt, 1);
		htab_init(&th->rt, d->nkeys, d->naggrs + 1, 16, 1, 0, vs);
		htab_build_index_local(&th->rt, 1);
		for (int v = 0; v < q.ty.nv; v++)
			th->v[v] = (int64_t*)xalloc(vs * 8);
		for (int n = 0; n < d->nexpr; n++)
			th->node[n] = (i128*)xalloc(vs * 16);
		for (int j = 0; j < d->njoins; j++) {
			th->lvl[j].bucket = (uint64_t*)xalloc(vs * 8);
			th->lvl[j].hash = (uint64_t*)xalloc(vs * 8);
			th->lvl[j].active = (sel_t*)xalloc(vs * 4);
			th->lvl[j].active2 = (sel_t*)xalloc(vs * 4);
			th->lvl[j].hit = (sel_t*)xalloc(vs * 4);
			th->lvl[j].match = (uint8_t*)xalloc(vs);
		}
		th->sel_a = (sel_t*)xalloc(vs * 4);
		th->sel_b = (sel_t*)xalloc(vs * 4);
		th->pred = (uint8_t*)xalloc(vs);
		th->g_hash = (uint64_t*)xalloc(vs * 8);
		th->g_bucket = (uint64_t*)xalloc(vs * 8);
		th->g_newpos = (uint64_t*)xalloc(vs * 8);
		th->g_miss = (sel_t*)xalloc(vs * 4);
		th->g_miss2 = (sel_t*)xalloc(vs * 4);
		th->g_hit = (sel_t*)xalloc(vs * 4);
		th->g_hit2 = (sel_t*)xalloc(vs * 4);
		th->g_found = (sel_t*)xalloc(vs * 4);
		th->g_tmp = (sel_t*)xalloc(vs * 4);
		th->g_pred = (uint8_t*)xalloc(vs);
		th->g_eq = (uint8_t*)xalloc(vs);
		for (int k = 0; k < VH_MAX_KEYS; k++)
			th->keyv[k] = (int64_t*)xalloc(vs * 8);
		for (int a = 0; a < VH_MAX_AGGRS; a++)
			th->aggv[a] = (i128*)xalloc(vs * 16);
	}
	struct timespec t0, t1;
	clock_gettime(CLOCK_MONOTONIC, &t0);
	pthread_t* tids = (pthread_t*)xalloc(sizeof(pthread_t) * (size_t)nthreads);
	for (int t = 1; t < nthreads; t++)
		pthread_create(&tids[t], NULL, worker, &q.th[t]);
	worker(&q.th[0]);
	for (int t = 1; t < nthreads; t++)
		pthread_join(tids[t], NULL);
	clock_gettime(CLOCK_MONOTONIC, &t1);
	if (elapsed_ns)
		*elapsed_ns = (int64_t)(t1.tv_sec - t0.tv_sec) * 1000000000LL + (t1.tv_nsec - t0.tv_nsec);

	size_t total = 0;
	int64_t rows = 0;
	for (int t = 0; t < nthreads; t++) {
		total += q.th[t].out_len;
		rows += (int64_t)q.th[t].out_rows;
	}
	*len = total;
	if (nrows)
		*nrows = rows;
	if (out && cap > total) {
		size_t o = 0;
		for (int t = 0; t < nthreads; t++) {
			if (q.th[t].out_len)
				memcpy(out + o, q.th[t].out, q.th[t].out_len);
			o += q.th[t].out_len;
		}
		out[o] = 0;
	} else if (out) {
		rc = VH_ERR_CAPACITY;
	}

	for (int t = 0; t < nthreads; t++) {
		vo_thread* th = &q.th[t];
		htab_free(&th->gt);
		htab_free(&th->rt);
		for (int v = 0; v < q.ty.nv; v++) free(th->v[v]);
		for (int n = 0; n < d->nexpr; n++) free(th->node[n]);
		for (int j = 0; j < d->njoins; j++) {
			free(th->lvl[j].bucket); free(th->lvl[j].hash); free(th->lvl[j].active);
			free(th->lvl[j].active2); free(th->lvl[j].hit); free(th->lvl[j].match);
		}
		free(th->sel_a); free(th->sel_b); free(th->pred);
		free(th->g_hash); free(th->g_bucket); free(th->g_newpos); free(th->g_miss);
		free(th->g_miss2); free(th->g_hit); free(th->g_hit2); free(th->g_found);
		free(th->g_tmp); free(th->g_pred); free(th->g_eq);
		for (int k = 0; k < VH_MAX_KEYS; k++) free(th->keyv[k]);
		for (int a = 0; a < VH_MAX_AGGRS; a++) free(th->aggv[a]);
		free(th->out);
	}
	for (int j = 0; j < d->njoins; j++)
		htab_free(&q.jt[j]);
	free(q.th);
	free(tids);
	pthread_barrier_destroy(&q.barrier);
	return rc;
}

/* ======================================================================== */
/* Closed-form checker                                                       */
/* ======================================================================== */

typedef struct bf_entry { int64_t key[VH_MAX_JOIN_KEYS]; int64_t row; } bf_entry;

static int bf_key_cmp(const int64_t* x, const int64_t* y)
{
	for (int k = 0; k < VH_MAX_JOIN_KEYS; k++)
		if (x[k] != y[k])
			return x[k] < y[k] ? -1 : 1;
	return 0;
}

static int bf_entry_cmp(const void* a, const void* b)
{
	const bf_entry* x = (const bf_entry*)a;
	const bf_entry* y = (const bf_entry*)b;
	const int c = bf_key_cmp(x->key, y->key);
	if (c)
		return c;
	return x->row < y->row ? -1 : x->row > y->row;
}

typedef struct bf_ctx {
	const vh_pipeline_desc* d;
	vo_typing ty;
	bf_entry* sorted[VH_MAX_JOINS];
	int64_t nsorted[VH_MAX_JOINS];
	int64_t vals[VH_MAX_COLS + VH_MAX_JOINS * VH_MAX_PAYLOAD];
	int rec_words;          /* keys + aggregates, as i128 words        */
	i128* recs;
	int64_t nrecs, caprecs;
} bf_ctx;

static void bf_emit(bf_ctx* c)
{
	const vh_pipeline_desc* d = c->d;
	if (c->nrecs == c->caprecs) {
		c->caprecs = c->caprecs ? c->caprecs * 2 : 1 << 16;
		c->recs = (i128*)realloc(c->recs, (size_t)c->caprecs * (size_t)c->rec_words * 16);
		if (!c->recs) { fprintf(stderr, "oracle: out of memory\n"); abort(); }
	}
	i128* r = c->recs + c->nrecs * c->rec_words;
	i128 node[VH_MAX_EXPR];
	for (int n = 0; n < d->nexpr; n++) {
		const vh_expr* e = &d->expr[n];
		switch (e->op) {
		case VH_E_COL: node[n] = widen(c->ty.vcol[e->a].type, c->vals[e->a]); break;
		case VH_E_CONST: node[n] = e->value; break;
		case VH_E_ADD: node[n] = node[e->a] + node[e->b]; break;
		case VH_E_SUB: node[n] = node[e->a] - node[e->b]; break;
		default: node[n] = node[e->a] * node[e->b]; break;
		}
	}
	for (int k = 0; k < d->nkeys; k++)
		r[k] = widen(c->ty.vcol[d->keys[k]].type, c->vals[d->keys[k]]);
	for (int a = 0; a < d->naggrs; a++)
		r[d->nkeys + a] = d->aggrs[a].fn == VH_A_COUNT ? 1 : node[d->aggrs[a].expr];
	c->nrecs++;
}

static void bf_join(bf_ctx* c, int j, int vbase)
{
	const vh_pipeline_desc* d = c->d;
	if (j == d->njoins) {
		bf_emit(c);
		return;
	}
	const vh_join_desc* jd = &d->joins[j];
	int64_t key[VH_MAX_JOIN_KEYS] = {0};
	for (int k = 0; k < join_nkeys(jd); k++)
		key[k] = c->vals[join_probe_key(jd, k)];
	/* lower bound */
	int64_t lo = 0, hi = c->nsorted[j];
	while (lo < hi) {
		int64_t mid = (lo + hi) / 2;
		if (bf_key_cmp(c->sorted[j][mid].key, key) < 0) lo = mid + 1; else hi = mid;
	}
	for (int64_t m = lo; m < c->nsorted[j] && bf_key_cmp(c->sorted[j][m].key, key) == 0; m++) {
		int64_t row = c->sorted[j][m].row;
		for (int p = 0; p < jd->npayload; p++)
			c->vals[vbase + p] = load_col(&jd->build.cols[jd->payload[p]], row);
		bf_join(c, j + 1, vbase + jd->npayload);
		if (jd->variant == VH_JOIN01)
			break;
	}
}

static int g_bf_nkeys;
static int bf_rec_cmp(const void* a, const void* b)
{
	const i128* x = (const i128*)a;
	const i128* y = (const i128*)b;
	for (int k = 0; k < g_bf_nkeys; k++)
		if (x[k] != y[k])
			return x[k] < y[k] ? -1 : 1;
	return 0;
}

/* Exact group-by: every qualifying (joined) tuple becomes one record of
 * i128 words; records are sorted by key and reduced.  Only groups that
 * received a tuple exist (relalg_translator.cpp:698-701). */
int vo_bruteforce_pipeline(const vh_pipeline_desc* d, char* out, size_t cap, size_t* len,
	int64_t* nrows)
{
	bf_ctx c;
	memset(&c, 0, sizeof(c));
	c.d = d;
	int rc = vo_type_pipeline(d, &c.ty);
	if (rc)
		return rc;
	c.rec_words = d->nkeys + d->naggrs;
	if (c.rec_words == 0)
		c.rec_words = 1;
	for (int j = 0; j < d->njoins; j++) {
		const vh_join_desc* jd = &d->joins[j];
		c.sorted[j] = (bf_entry*)xalloc(sizeof(bf_entry) * (size_t)(jd->build.rows + 1));
		int64_t n = 0;
		for (int64_t r = 0; r < jd->build.rows; r++) {
			if (!probe_pass(&jd->build, r))
				continue;
			memset(c.sorted[j][n].key, 0, sizeof(c.sorted[j][n].key));
			for (int k = 0; k < join_nkeys(jd); k++)
				c.sorted[j][n].key[k] = load_col(&jd->build.cols[join_build_key(jd, k)], r);
			c.sorted[j][n].row = r;
			n++;
		}
		qsort(c.sorted[j], (size_t)n, sizeof(bf_entry), bf_entry_cmp);
		c.nsorted[j] = n;
	}
	for (int64_t r = 0; r < d->probe.rows; r++) {
		if (!probe_pass(&d->probe, r))
			continue;
		for (int col = 0; col < d->probe.ncols; col++)
			c.vals[col] = load_col(&d->probe.cols[col], r);
		bf_join(&c, 0, d->probe.ncols);
	}
	g_bf_nkeys = d->nkeys;
	if (c.nrecs)
		qsort(c.recs, (size_t)c.nrecs, (size_t)c.rec_words * 16, bf_rec_cmp);

	vo_thread th;
	memset(&th, 0, sizeof(th));
	int64_t groups = 0;
	for (int64_t i = 0; i < c.nrecs;) {
		int64_t e = i + 1;
		while (e < c.nrecs && bf_rec_cmp(c.recs + i * c.rec_words, c.recs + e * c.rec_words) == 0)
			e++;
		i128* first = c.recs + i * c.rec_words;
		for (int k = 0; k < d->nkeys; k++)
			out_value(&th, k == 0, first[k]);
		for (int a = 0; a < d->naggrs; a++) {
			i128 s = 0;
			for (int64_t m = i; m < e; m++)
				s += c.recs[m * c.rec_words + d->nkeys + a];
			out_value(&th, d->nkeys == 0 && a == 0, s);
		}
		out_push(&th, "\n", 1);
		groups++;
		i = e;
	}
	*len = th.out_len;
	if (nrows)
		*nrows = groups;
	if (out && cap > th.out_len) {
		if (th.out_len)
			memcpy(out, th.out, th.out_len);
		out[th.out_len] = 0;
	} else if (out) {
		rc = VH_ERR_CAPACITY;
	}
	free(th.out);
	free(c.recs);
	for (int j = 0; j < d->njoins; j++)
		free(c.sorted[j]);
	return rc;
}

size_t vo_sizeof_pipeline_desc(void)
{
	return sizeof(vh_pipeline_desc);
}
