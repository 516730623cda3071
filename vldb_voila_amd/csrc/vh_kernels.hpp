// Host-side declarations of the kernel launchers in vh_pipeline.hip.
#pragma once

#include <hip/hip_runtime.h>

#include "vh_plan.hpp"

namespace vh {

// p[0, n) = v; with only_if, only when *only_if != 0 (read on the device)
hipError_t launch_fill(u64* p, u64 n, u64 v, hipStream_t s, const u32* only_if = nullptr);
hipError_t launch_build(const BuildPlan* dB, int grid, hipStream_t s);
// Partitioned build of a direct table (B.nparts > 0): per-chunk LDS sort,
// then one LDS-built slice per partition.  B: host copy; dB: device copy.
hipError_t launch_part_build(const BuildPlan& B, const BuildPlan* dB, hipStream_t s);
// The same two steps over batch work items: scatter items (BatchItem::build,
// chunk r0) or, with `slices`, partition items (build, partition r0), taken
// from a queue by `grid` persistent workgroups; slice_bytes = the largest
// part_slice_lds_bytes of the builds involved.
hipError_t launch_batch_part(bool slices, const BuildPlan* d_builds, const BatchItem* d_items, u32 first, u32 count,
	u32* d_counter, int grid, size_t slice_bytes, hipStream_t s);
size_t part_slice_lds_bytes(const BuildPlan& B);
static constexpr int kMaxBuildParts = 4096;      // partitions of one build
static constexpr int kMaxBuildChunks = 2048;     // 4096-row chunks of one build
static constexpr int kBuildChunkRows = 4096;
static constexpr int kMaxSliceBytes = 32768;     // LDS slice of a partition
// P: host copy (launch shape); dP: the same plan in device memory.
// launch_pipeline runs the generic fused kernel; GB_REG / GB_LDS then need
// launch_merge.
hipError_t launch_pipeline(const Plan& P, const Plan* dP, int grid, hipStream_t s);
// The generic kernels of one group-by strategy (vh_modes.hip, one translation
// unit per strategy): the pipeline kernel with `lds` bytes of dynamic LDS,
// and the batched probe kernel over work items (every strategy but GB_REG).
template <int MODE>
hipError_t launch_pipeline_mode(const Plan* dP, int grid, size_t lds, hipStream_t s);
template <int MODE>
hipError_t launch_batch_probe_mode(const Plan* d_plans, const BatchItem* d_items, u32 first, u32 count,
	u32* d_counter, int grid, size_t lds_bytes, hipStream_t s);
#define VH_DECLARE_MODE(M) \
	template <> hipError_t launch_pipeline_mode<M>(const Plan*, int, size_t, hipStream_t);
VH_DECLARE_MODE(GB_REG)
VH_DECLARE_MODE(GB_LDS)
VH_DECLARE_MODE(GB_GLOBAL)
VH_DECLARE_MODE(GB_HASH)
VH_DECLARE_MODE(GB_LDSHASH)
VH_DECLARE_MODE(GB_PART)
#undef VH_DECLARE_MODE
#define VH_DECLARE_MODE(M) \
	template <> hipError_t launch_batch_probe_mode<M>(const Plan*, const BatchItem*, u32, u32, u32*, int, size_t, \
		hipStream_t);
VH_DECLARE_MODE(GB_LDS)
VH_DECLARE_MODE(GB_GLOBAL)
VH_DECLARE_MODE(GB_HASH)
VH_DECLARE_MODE(GB_LDSHASH)
VH_DECLARE_MODE(GB_PART)
#undef VH_DECLARE_MODE
// GB_REG / GB_LDS: merge the per-block partials of a `grid`-block launch
hipError_t launch_merge(const Plan& P, const Plan* dP, int grid, hipStream_t s);
// Dynamic LDS of the probe kernel: the group-by state, then the join tables
// kept in LDS (JoinT::lds)
size_t pipeline_lds_bytes(const Plan& P);
// LDS of the group-by state alone (where the join tables start, rounded up)
size_t groupby_lds_bytes(const Plan& P);
// Bytes of join j's LDS copy (JoinT::lds), rounded up to 16
size_t join_lds_bytes(const JoinT& J);
// bits[k / 32] bit k % 32 = (dir entry k != all-ones), k < nkeys
hipError_t launch_dir_bits(const void* dir, int w, u64 nkeys, u32* bits, hipStream_t s);
// GB_PART: LDS of part_agg_kernel's slice (2^gshift groups)
size_t part_slice_bytes(const Plan& P);

} // namespace vh
