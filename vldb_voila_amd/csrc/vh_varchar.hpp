// Launcher of the varchar statement kernels (vh_varchar.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace vh {

enum VcOp { VC_HASH = 0, VC_CONTAINS = 1, VC_EQ = 2, VC_NE = 3 };
static constexpr int kVcMaxLen = 255;     // u8 length indicator
static constexpr int kVcMaxConst = 256;   // needle / constant bytes, at most kVcMaxLen used

// op VC_HASH: res u64[], h = previous hashes (rehash) or nullptr;
// VC_CONTAINS: res u8[], needle (cbytes, clen);
// VC_EQ / VC_NE: res u8[], against col2 (max_len2) or, when col2 is null,
// the constant (cbytes, clen).  cbytes is host memory.
hipError_t launch_varchar(int op, const int32_t* sel, long long inum, void* res, const uint8_t* col,
	int max_len, const unsigned long long* h, const uint8_t* col2, int max_len2, const uint8_t* cbytes,
	unsigned clen, hipStream_t s);

} // namespace vh
