// Generated probe kernels (vh_jit.cpp): HIP source for one lowered plan,
// compiled with hipRTC and cached per (gfx target, source text).
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "vh_plan.hpp"

namespace vh {

// Rows per lane of the generated kernel a configuration asks for; 0 = the
// generator's default (jit_source).
inline int jit_unroll(const vh_config* cfg)
{
	return (cfg && cfg->unroll >= 1 && cfg->unroll <= 16) ? cfg->unroll : 0;
}

// Source of the fused probe kernel for P (its mode, joins, keys and sums).
// unroll = rows per lane per chunk (0: 8 for the vector loop when a probe
// row is at most 32 bytes of columns, else 2; 4 for the scalar loop).  eager: every probe column is loaded
// before the predicates; then, when every column is aligned to a pair of its
// values and unroll is even, full chunks of 64 * unroll rows go through the
// vector loop (2-row vector loads per lane, contiguous per instruction) and
// only the last partial chunk through the scalar loop.  Not eager: scalar
// loop, columns no predicate reads are loaded for qualifying rows only.
// pipeline: the vector loop requests the next chunk's columns right behind
// the last join's table lookups.
// block: threads per workgroup (the kernel's launch bound).
// batch: the entry point of a batched launch, vh_generated_batch(plans,
// items, first, count, counter), which takes work items (rows [r0, r1) of
// plans[plan]) from the queue until it is empty and adds each item's LDS
// group-by state to the pipeline's global arrays.
std::string jit_source(const Plan& P, bool eager, int unroll, bool pipeline, int block = 256, bool batch = false);

// Generated-kernel options of a configuration's vh_kernel value.
inline bool jit_eager(int kernel) { return kernel != VH_KERNEL_JIT_LAZY; }
inline bool jit_pipeline(int kernel) { return kernel != VH_KERNEL_JIT_LAZY; }
// Compile (or fetch from the cache) `src` for the current device.
// `name`: the kernel's symbol (vh_generated or vh_generated_batch).
int jit_compile(const std::string& src, hipFunction_t* fn, std::string& log, const char* name = "vh_generated");
// Launch a compiled kernel over P (dP: its device copy).  GB_REG / GB_LDS
// leave per-block partials for launch_merge.
// start / stop (may be null): events the runtime stamps with the kernel's
// own start and end (hipExtModuleLaunchKernel).
// lds_bytes: dynamic LDS of the launch (< 0: pipeline_lds_bytes(P)).
hipError_t launch_generated(hipFunction_t fn, const Plan& P, const Plan* dP, int grid, int block, hipStream_t s,
	hipEvent_t start = nullptr, hipEvent_t stop = nullptr, long lds_bytes = -1);
// Radix-partitioned probe (P.radix): jit_source gives the probe kernel that
// writes the partitioned tuples; this is the join kernel that consumes them
// (vh_generated_join, one workgroup per partition of the last join).
std::string jit_join_source(const Plan& P, int unroll, int block);
// The values a radix probe tuple of P carries: their widths in bytes
// (P.rval[k] holds value k), and the tuple's bytes with its key offset.
int jit_radix_values(const Plan& P, std::vector<int>& widths);
// LDS bytes the probe kernel stages one chunk of `chunk` rows' tuples in
// (after P.rlds_a; P.rparts must be set).
size_t jit_radix_stage_bytes(const Plan& P, u64 chunk);

} // namespace vh
