// Learned flavor selector: FP32 MLP 6 -> 64 -> 64 -> 48 (ReLU), softmax over
// the kernel variants, trained with softmax cross-entropy and Adam.  It
// replaces VOILA's exhaustive flavor exploration (explorer.cpp:112-427) as
// the first guess; the autotuner's exhaustive timing remains the fallback.
//
// Every reduction runs in a fixed order (one thread per output element,
// looping over the batch), so training is deterministic run to run.
#include <hip/hip_runtime.h>

#include <stdint.h>

#include "vh_mlp.hpp"

namespace vh {

// parameter layout (floats): W1[IN][H] b1[H] W2[H][H] b2[H] W3[H][OUT] b3[OUT]
__host__ __device__ constexpr int off_w1() { return 0; }
__host__ __device__ constexpr int off_b1() { return kIn * kHid; }
__host__ __device__ constexpr int off_w2() { return off_b1() + kHid; }
__host__ __device__ constexpr int off_b2() { return off_w2() + kHid * kHid; }
__host__ __device__ constexpr int off_w3() { return off_b2() + kHid; }
__host__ __device__ constexpr int off_b3() { return off_w3() + kHid * kOut; }
static_assert(off_b3() + kOut == kParams, "parameter layout");

// h1 = relu(x W1 + b1): one thread per (sample, unit)
__global__ void mlp_layer1(const float* __restrict__ P, const float* __restrict__ X, float* __restrict__ H1, int B)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= B * kHid)
		return;
	const int b = i / kHid, j = i % kHid;
	float a = P[off_b1() + j];
	for (int k = 0; k < kIn; k++)
		a += X[b * kIn + k] * P[off_w1() + k * kHid + j];
	H1[i] = a > 0.f ? a : 0.f;
}

__global__ void mlp_layer2(const float* __restrict__ P, const float* __restrict__ H1, float* __restrict__ H2, int B)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= B * kHid)
		return;
	const int b = i / kHid, j = i % kHid;
	float a = P[off_b2() + j];
	for (int k = 0; k < kHid; k++)
		a += H1[b * kHid + k] * P[off_w2() + k * kHid + j];
	H2[i] = a > 0.f ? a : 0.f;
}

// logits = h2 W3 + b3, masked softmax per sample (mask 0 = variant invalid)
__global__ void mlp_layer3_softmax(const float* __restrict__ P, const float* __restrict__ H2,
	const uint8_t* __restrict__ mask, float* __restrict__ prob, int B)
{
	const int b = blockIdx.x * blockDim.x + threadIdx.x;
	if (b >= B)
		return;
	float z[kOut];
	float mx = -3.0e38f;
	for (int o = 0; o < kOut; o++) {
		float a = P[off_b3() + o];
		for (int k = 0; k < kHid; k++)
			a += H2[b * kHid + k] * P[off_w3() + k * kOut + o];
		z[o] = a;
		if ((!mask || mask[b * kOut + o]) && a > mx)
			mx = a;
	}
	float s = 0.f;
	for (int o = 0; o < kOut; o++) {
		const float e = (!mask || mask[b * kOut + o]) ? __expf(z[o] - mx) : 0.f;
		z[o] = e;
		s += e;
	}
	for (int o = 0; o < kOut; o++)
		prob[b * kOut + o] = s > 0.f ? z[o] / s : 0.f;
}

// dlogits = (prob - onehot(y)) / B ; loss terms -log p[y]
__global__ void mlp_dlogits(const float* __restrict__ prob, const int32_t* __restrict__ y, float* __restrict__ dz,
	float* __restrict__ loss_terms, int B)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= B * kOut)
		return;
	const int b = i / kOut, o = i % kOut;
	const float p = prob[i];
	dz[i] = (p - (o == y[b] ? 1.f : 0.f)) / (float)B;
	if (o == y[b])
		loss_terms[b] = -__logf(p > 1e-30f ? p : 1e-30f);
}

// dA = dZ W^T masked by relu(A) > 0, for a layer with `n_out` outputs
__global__ void mlp_back_act(const float* __restrict__ W, int n_out, const float* __restrict__ dZ,
	const float* __restrict__ A, float* __restrict__ dA, int B)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= B * kHid)
		return;
	const int b = i / kHid, k = i % kHid;
	float s = 0.f;
	for (int o = 0; o < n_out; o++)
		s += dZ[b * n_out + o] * W[k * n_out + o];
	dA[i] = A[i] > 0.f ? s : 0.f;
}

// dW[k][o] = sum_b In[b][k] dZ[b][o]; db[o] = sum_b dZ[b][o]
__global__ void mlp_back_weights(const float* __restrict__ In, int n_in, const float* __restrict__ dZ, int n_out,
	float* __restrict__ dW, float* __restrict__ db, int B)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n_in * n_out) {
		const int k = i / n_out, o = i % n_out;
		float s = 0.f;
		for (int b = 0; b < B; b++)
			s += In[b * n_in + k] * dZ[b * n_out + o];
		dW[i] = s;
	} else if (i < n_in * n_out + n_out) {
		const int o = i - n_in * n_out;
		float s = 0.f;
		for (int b = 0; b < B; b++)
			s += dZ[b * n_out + o];
		db[o] = s;
	}
}

__global__ void mlp_adam(float* __restrict__ P, const float* __restrict__ G, float* __restrict__ M,
	float* __restrict__ V, int n, float lr, float b1, float b2, float eps, float c1, float c2)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	const float g = G[i];
	const float m = b1 * M[i] + (1.f - b1) * g;
	const float v = b2 * V[i] + (1.f - b2) * g * g;
	M[i] = m;
	V[i] = v;
	const float mh = m / c1, vh = v / c2;
	P[i] -= lr * mh / (sqrtf(vh) + eps);
}

static unsigned blocks(int n) { return (unsigned)((n + 255) / 256); }

hipError_t mlp_forward(const float* P, const float* X, const uint8_t* mask, float* H1, float* H2, float* prob, int B,
	hipStream_t s)
{
	hipLaunchKernelGGL(mlp_layer1, dim3(blocks(B * kHid)), dim3(256), 0, s, P, X, H1, B);
	hipLaunchKernelGGL(mlp_layer2, dim3(blocks(B * kHid)), dim3(256), 0, s, P, H1, H2, B);
	hipLaunchKernelGGL(mlp_layer3_softmax, dim3(blocks(B)), dim3(256), 0, s, P, H2, mask, prob, B);
	return hipGetLastError();
}

hipError_t mlp_backward(const float* P, const float* X, const float* H1, const float* H2, const float* prob,
	const int32_t* y, float* G, float* dz, float* dh2, float* dh1, float* loss_terms, int B, hipStream_t s)
{
	hipLaunchKernelGGL(mlp_dlogits, dim3(blocks(B * kOut)), dim3(256), 0, s, prob, y, dz, loss_terms, B);
	hipLaunchKernelGGL(mlp_back_weights, dim3(blocks(kHid * kOut + kOut)), dim3(256), 0, s, H2, kHid, dz, kOut,
		G + off_w3(), G + off_b3(), B);
	hipLaunchKernelGGL(mlp_back_act, dim3(blocks(B * kHid)), dim3(256), 0, s, P + off_w3(), kOut, dz, H2, dh2, B);
	hipLaunchKernelGGL(mlp_back_weights, dim3(blocks(kHid * kHid + kHid)), dim3(256), 0, s, H1, kHid, dh2, kHid,
		G + off_w2(), G + off_b2(), B);
	hipLaunchKernelGGL(mlp_back_act, dim3(blocks(B * kHid)), dim3(256), 0, s, P + off_w2(), kHid, dh2, H1, dh1, B);
	hipLaunchKernelGGL(mlp_back_weights, dim3(blocks(kIn * kHid + kHid)), dim3(256), 0, s, X, kIn, dh1, kHid,
		G + off_w1(), G + off_b1(), B);
	return hipGetLastError();
}

hipError_t mlp_adam_step(float* P, const float* G, float* M, float* V, float lr, int t, hipStream_t s)
{
	const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
	float c1 = 1.f, c2 = 1.f;
	for (int i = 0; i < t; i++) {
		c1 *= b1;
		c2 *= b2;
	}
	hipLaunchKernelGGL(mlp_adam, dim3(blocks(kParams)), dim3(256), 0, s, P, G, M, V, kParams, lr, b1, b2, eps,
		1.f - c1, 1.f - c2);
	return hipGetLastError();
}

} // namespace vh
