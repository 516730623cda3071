// Learned flavor selector MLP: dimensions and kernel launchers (vh_mlp.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace vh {

static constexpr int kIn = 6;      // log rows, log build rows, selectivity, log groups, theta, key width
static constexpr int kHid = 64;
static constexpr int kOut = 48;    // kernel variants
static constexpr int kParams = kIn * kHid + kHid + kHid * kHid + kHid + kHid * kOut + kOut;   // 7,728

hipError_t mlp_forward(const float* P, const float* X, const uint8_t* mask, float* H1, float* H2, float* prob, int B,
	hipStream_t s);
hipError_t mlp_backward(const float* P, const float* X, const float* H1, const float* H2, const float* prob,
	const int32_t* y, float* G, float* dz, float* dh2, float* dh1, float* loss_terms, int B, hipStream_t s);
hipError_t mlp_adam_step(float* P, const float* G, float* M, float* V, float lr, int t, hipStream_t s);

} // namespace vh
