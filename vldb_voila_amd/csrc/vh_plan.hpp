// Lowered form of a vh_pipeline_desc, passed by value to the fused gfx950
// kernels.  Every field is uniform across the grid, so the kernels branch on
// it with scalar instructions only.
#pragma once

#ifndef __HIPCC_RTC__   // hipRTC provides these itself
#include <stdint.h>
#endif

#include "../../include/voila_hip.h"

namespace vh {

typedef unsigned long long u64;
typedef long long i64;
typedef unsigned int u32;

// Where a value of the probe stream comes from.
enum RoleKind { ROLE_NONE = 0, ROLE_COL = 1, ROLE_PAY = 2, ROLE_CONST = 3 };

struct Role {
	const void* ptr;     // ROLE_COL: probe column base address
	i64 c;               // ROLE_CONST: value
	int kind;
	int type;            // ROLE_COL: enum vh_type of the column
	int slot;            // ROLE_PAY: join index
	int item;            // ROLE_PAY: which of the join's payload columns
};

// A linear term sa*a + sb*b + c with sa, sb in {-1, 0, 1}.  A sum argument is
// the product of up to three terms, e.g. price * (100 - disc).
struct Term {
	Role a;
	Role b;
	i64 c;
	int sa;
	int sb;
};

struct SumArg {
	Term t[3];
	int nterms;
	int wide;            // result i128 and argument wider than 32 bits: lanes
	                     // track a carry word; narrower arguments cannot
	                     // overflow a 64-bit lane sum
	int arg_signed;      // argument type signed (sign extends into i128)
	int res128;          // result typed i128 (global modes accumulate exactly)
	int narrow;          // the argument's exact interval fits 32 bits (signed
	                     // if arg_signed): a partitioned group-by's tuples
	                     // keep it in 4 bytes
	int toff;            // its byte offset in such a tuple
};

struct Pred {
	Role r;
	i64 value;
	int cmp;
	int pad;
};

// Join table.  Hash layout: open addressing over 64-bit slots, low word =
// key - kmin, high word = payload - pmin (both < 2^32 - 1), all-ones = empty,
// linear probing from voila_hash(key - kmin) & mask.  The probe sequence
// wraps within an aligned block of smask + 1 slots: the whole table
// (smask == mask), or, for a table built partitioned, the sub-table of the
// key's partition (see BuildPlan); rows that found their sub-table full are
// in the overflow table `ovf` (capacity mask + 1, whole-table wrap), used
// only when the build set `spill`.  Direct layout (dir_w != 0):
// dir[key - kmin] holds payload - pmin in dir_w bytes (all-ones = no row);
// a key's second and later rows go to the hash slots, which then serve only
// as its overflow (dups set).
struct JoinT {
	u64* slots;
	u64 mask;
	u64 smask;
	u64* ovf;
	const u32* spill;
	const void* dir;
	int dir_w;
	int hashed;          // partitioned hash table (smask < mask, ovf and spill set)
	Role key;            // probe-side key (nkeys == 1)
	i64 kmin;            // key range; 0 .. the largest packed key when nkeys > 1
	i64 kmax;
	i64 pmin;            // payload 0's minimum (npay == 1)
	int npay;
	// multi-column keys (nkeys > 1): the entry's key offset is the sum of
	// (column value - kminc[c]) << kshiftc[c]; a value outside
	// [kminc[c], kminc[c] + krangec[c]] matches no row
	int nkeys;
	Role keyc[VH_MAX_JOIN_KEYS];
	i64 kminc[VH_MAX_JOIN_KEYS];
	u64 krangec[VH_MAX_JOIN_KEYS];
	int kshiftc[VH_MAX_JOIN_KEYS];
	int kpad;
	// payload columns packed in the entry's payload word: payload k is
	// ((word >> pshiftc[k]) & pmaskc[k]) + pminc[k]
	i64 pminc[VH_MAX_PAYLOAD];
	int pshiftc[VH_MAX_PAYLOAD];
	u32 pmaskc[VH_MAX_PAYLOAD];
	int variant;         // enum vh_join_variant
	int last;            // last join of the pipeline: JoinN may match many
	int filtered;        // the build side has predicates (probe keys often miss)
	const u32* dups;     // set by the build when a key occurs twice
	// generated probe kernels: a copy of the table in LDS at byte lds_off of
	// the dynamic LDS, made by every workgroup at its start.  LDS_DIR: the
	// whole dir[]; LDS_BITS: one bit per key of [kmin, kmax] (set iff dir
	// holds a row), from `bits`, rewritten after every build by
	// dir_bits_kernel, so that the lookups of keys without a row stay in LDS
	int lds;
	u32 lds_off;
	const u32* bits;
};

enum JoinLds { LDS_NONE = 0, LDS_DIR = 1, LDS_BITS = 2 };

// Bits of a probe launch's join flags word: bit j = join j's build saw a
// repeated key, bit kSpillBit + j = join j's partitioned hash build spilled
// rows to its overflow table.
static constexpr int kSpillBit = 16;

struct KeyT {
	Role r;
	i64 min;
	u64 mult;            // dense index multiplier (direct modes)
	u64 range;           // max - min
	int shift;           // bit offset of (v - min) in the packed key (hash mode)
	int pad;
};

// values match enum vh_group_mode (5 there is VH_GROUP_DENSE, which the
// lowering resolves to GB_REG or GB_LDS)
enum GroupMode { GB_REG = 0, GB_LDS = 1, GB_GLOBAL = 2, GB_HASH = 3, GB_LDSHASH = 4, GB_PART = 6 };

// GB_PART: partitioned dense group-by for domains too large for LDS.  The
// probe rows are cut into regions of gregion_rows rows (a workgroup's slab,
// or a batched work item).  A region keeps one tuple per qualifying row: a
// record of grec bytes (a multiple of 8) holding the group index as a u32 at
// byte 0 and each sum's argument at SumArg::toff (4 bytes when
// SumArg::narrow, else 8, 8-aligned), counted per partition (g >> gshift) in
// LDS.  The region's first gstage tuples are staged in LDS, the rest go to
// gtup0 at their rows' own positions; at its end the region stores them
// ordered by partition into gtup1 (its rows' positions) and records, per
// partition, the run's end (gend, inclusive prefix of the counts) in a
// [region][partition] table.
// part_agg_kernel then aggregates each partition's runs in an LDS slice and
// adds the slice into out_*.  A row with more than one match (JoinN) that
// finds its region's tuple space full adds straight into out_* with atomics.
static constexpr int kMaxGroupParts = 4096;
static constexpr int kMaxGroupRegions = 2048;
static constexpr int kPartStageBytes = 32 * 1024;   // LDS staging of a region's tuples

// GB_LDSHASH: per-block LDS open-addressing table over the dense group index,
// kLdsHashSlots entries, at most kLdsHashProbes probes; a tuple that finds no
// slot goes straight to the global arrays (same layout as GB_GLOBAL).
static constexpr int kLdsHashSlots = 1024;
static constexpr int kLdsHashProbes = 32;

static constexpr int kRegSlots = 8;    // dense groups held in registers
static constexpr int kRegSums = 4;     // sums held in registers

struct Plan {
	i64 rows;
	int npreds;
	int njoins;
	int nkeys;
	int nsums;
	int wide;            // some sum needs exact 128-bit bookkeeping (see SumArg)
	int mode;            // enum GroupMode
	int flush;           // GB_REG / GB_LDS: workgroups add into out_* with
	                     // atomics instead of leaving partials to merge
	int pad_flush;
	Pred preds[VH_MAX_PREDS];
	JoinT joins[VH_MAX_JOINS];
	KeyT keys[VH_MAX_KEYS];
	SumArg sums[VH_MAX_AGGRS];
	u64 domain;          // dense group domain (direct modes)
	// outputs: lo[s * groups + g], hi[...], cnt[g]; groups = domain or hash capacity
	u64* out_lo;
	i64* out_hi;
	u64* out_cnt;
	// per-block partials, element-major (GB_REG / GB_LDS), merged by merge_kernel
	u64* partials;
	// hash mode: packed keys, EMPTY = ~0
	u64* hkeys;
	u64 hmask;
	// partitioned mode (GB_PART)
	u64* gtup0;
	u64* gtup1;
	int grec;            // bytes of a tuple record
	int gstage;          // tuples of a region staged in LDS
	u32* gend;
	i64 gregion_rows;
	int gshift;
	int gparts;
	u32* error;          // bit 0: hash table full, bit 3: group key outside [min, max]
	// Radix-partitioned probe of the last join (radix != 0; that join's
	// table is a partitioned hash table, see BuildPlan).  The probe kernel
	// takes the rows in chunks of at most rchunk rows, one per workgroup
	// iteration, numbered from 0; chunk ids up to rchunks that no iteration
	// used get empty runs.  Each row that reaches the last join leaves a tuple in the workgroup's
	// LDS: its key offset (key - kmin) and the rw values the group-by needs
	// (probe columns, earlier joins' payloads).  At the chunk's end the
	// workgroup stores the chunk's tuples sorted by probe partition (the
	// aligned block of 2^rpshift slots of the key's hash, rparts of them)
	// into rkey / rval[k] at [c * rchunk, ...) and the partitions' running
	// ends into rend[c * rparts + p] (so partition p's run of chunk c is
	// [rend[.. p - 1], rend[.. p]) of the chunk).  The join kernel takes one
	// partition per workgroup: its 2^rpshift slots copied into LDS, its run
	// of every chunk, the lookups and the group-by.
	u32* rkey;
	void* rval[16];      // kRadixMaxVals
	u32* rend;
	i64 rchunk;
	int radix;
	int rw;              // values per tuple
	int rchunks;         // chunk ids the run table holds
	int rparts;          // probe partitions
	int rpshift;         // log2 slots of a probe partition (>= the build's sub-table)
	u32 rlds_a;          // probe kernel: LDS byte offset of its tuple staging
	u32 rlds_b;          // join kernel: LDS byte offset of the sub-table
	int rpad;
};

static constexpr int kRadixMaxVals = 16;        // values of a radix probe tuple (Plan::rval)
static constexpr int kRadixMaxChunks = 4096;    // chunk ids of a radix probe (one scan in LDS)

// A work item of a batched launch: rows [r0, r1) of one pipeline's build or
// probe side.
struct BatchItem {
	int plan;            // pipeline index
	int phase;           // 0 = build rows of builds[build], 1 = probe rows of plans[plan]
	int build;           // index into the build plan array (phase 0)
	int reserved;
	i64 r0;
	i64 r1;
};

struct BuildPlan {
	i64 rows;
	int npreds;
	int key_type;
	Pred preds[VH_MAX_PREDS];
	const void* key;
	const void* pay;     // payload 0 (npay == 1)
	int pay_type;
	int npay;
	i64 kmin;            // key range (nkeys == 1); packed keys start at 0
	i64 kmax;
	i64 pmin;
	// multi-column keys and payloads, packed as JoinT describes
	int nkeys;
	int pad2;
	const void* keyc[VH_MAX_JOIN_KEYS];
	int keyc_type[VH_MAX_JOIN_KEYS];
	int kshiftc[VH_MAX_JOIN_KEYS];
	i64 kminc[VH_MAX_JOIN_KEYS];
	u64 krangec[VH_MAX_JOIN_KEYS];
	const void* payc[VH_MAX_PAYLOAD];
	int payc_type[VH_MAX_PAYLOAD];
	int pshiftc[VH_MAX_PAYLOAD];
	i64 pminc[VH_MAX_PAYLOAD];
	u64* slots;
	u64 mask;
	void* dir;           // direct layout (see JoinT), dir_w bytes per key
	int dir_w;
	int pad0;
	u32* dups;
	u32* error;          // bit 2: key outside [kmin, kmax]
	// partitioned build of a direct table (nparts > 0): partition p holds
	// the keys with (key - kmin) >> pshift == p.  Chunk b (rows
	// [b * 4096, (b + 1) * 4096)) leaves its selected rows sorted by
	// partition in pbuf[b * 4096 ...] as (key - kmin) | payload << 32, the
	// run of partition p being pcnt[p * nchunks + b] rows from offset
	// poff[p * nchunks + b] (partition-major, so a partition's runs are
	// contiguous for the workgroup that builds it).
	u64* pbuf;
	u32* pcnt;
	u32* poff;
	// key bitmap of dir (JoinT::bits), written with each slice when set
	u32* bits;
	int nchunks;
	int nparts;
	int pshift;
	// partitioned hash table (hashed != 0): partition p is the aligned block
	// of 2^pshift slots a key's voila_hash(key - kmin) & mask falls in; the
	// slice launch builds each block as an LDS sub-table and stores all of
	// it; a row that finds its sub-table full goes to `ovf` with a global CAS
	// and sets *spill
	int hashed;
	u64* ovf;
	u32* spill;
};

} // namespace vh
