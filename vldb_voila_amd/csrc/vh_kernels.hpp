// Host-side declarations of the kernel launchers in vh_pipeline.hip.
#pragma once

#include <hip/hip_runtime.h>

#include "vh_plan.hpp"

namespace vh {

hipError_t launch_fill(u64* p, u64 n, u64 v, hipStream_t s);
hipError_t launch_build(const BuildPlan* dB, int grid, hipStream_t s);
// P: host copy (launch shape); dP: the same plan in device memory
hipError_t launch_pipeline(const Plan& P, const Plan* dP, int grid, hipStream_t s);
size_t pipeline_lds_bytes(const Plan& P);

} // namespace vh
