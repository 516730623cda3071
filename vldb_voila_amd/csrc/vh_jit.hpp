// Generated probe kernels (vh_jit.cpp): HIP source for one lowered plan,
// compiled with hipRTC and cached per (gfx target, source text).
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "vh_plan.hpp"

namespace vh {

static constexpr int kJitUnroll = 4;   // default rows per lane per chunk of generated kernels

// Rows per lane of the generated kernel a configuration asks for.
inline int jit_unroll(const vh_config* cfg)
{
	return (cfg && cfg->unroll >= 1 && cfg->unroll <= 16) ? cfg->unroll : kJitUnroll;
}

// Source of the fused probe kernel for P (its mode, joins, keys and sums).
// eager: load every probe column before the predicates (else columns not
// read by a predicate are loaded only for qualifying rows); unroll = rows
// per lane per chunk; pipeline: request the next chunk's columns while this
// chunk's join lookups are in flight (implies eager).
std::string jit_source(const Plan& P, bool eager, int unroll, bool pipeline);

// Generated-kernel options of a configuration's vh_kernel value.
inline bool jit_eager(int kernel) { return kernel != VH_KERNEL_JIT_LAZY; }
inline bool jit_pipeline(int kernel) { return kernel == VH_KERNEL_JIT_PIPE; }
// Compile (or fetch from the cache) `src` for the current device.
int jit_compile(const std::string& src, hipFunction_t* fn, std::string& log);
// Launch a compiled kernel over P (dP: its device copy).  GB_REG / GB_LDS
// leave per-block partials for launch_merge.
hipError_t launch_generated(hipFunction_t fn, const Plan& P, const Plan* dP, int grid, hipStream_t s);

} // namespace vh
