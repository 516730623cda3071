// extern "C" entry points of the autotuner (exhaustive variant timing) and
// the learned selector (MLP training and inference on the GPU).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <new>
#include <vector>

#include "../../include/voila_hip.h"
#include "vh_engine.hpp"
#include "vh_mlp.hpp"

using namespace vh;

#define HIP_CALL(x)                                                                  \
	do {                                                                             \
		hipError_t _e = (x);                                                         \
		if (_e != hipSuccess)                                                        \
			return fail(VH_ERR_HIP, "%s: %s", #x, hipGetErrorString(_e));           \
	} while (0)

// Variant index -> knobs: 4 group-by strategies x 3 join-table layouts x
// 4 probe grid sizes = 48 (the analog of the explorer's blend space,
// explorer_helper.cpp:6-17).
static const int kModes[4] = {VH_GROUP_DENSE, VH_GROUP_LDSHASH, VH_GROUP_PART, VH_GROUP_HASH};
// join table: {hash layout, hash layout at twice the slots, direct when dense}
static const int kTableShift[3] = {0, 1, 0};
static const int kTableLayout[3] = {VH_LAYOUT_HASH, VH_LAYOUT_HASH, VH_LAYOUT_AUTO};
static const int kGrids[4] = {256, 512, 1024, 2048};

struct vh_mlp {
	float* P = nullptr;   // params
	float* G = nullptr;   // grads
	float* M = nullptr;
	float* V = nullptr;
	float* act = nullptr; // H1, H2, prob, dz, dh2, dh1, loss per sample
	int cap = 0;          // batch capacity
	int step = 0;
	~vh_mlp()
	{
		for (float* p : {P, G, M, V, act})
			if (p)
				(void)hipFree(p);
	}
};

extern "C" {

int vh_variant_count(void) { return 48; }

int vh_variant_config(int32_t v, vh_config* out)
{
	if (!out || v < 0 || v >= 48)
		return fail(VH_ERR_INVALID, "variant %d", v);
	memset(out, 0, sizeof(*out));
	out->group_mode = kModes[v / 12];
	out->table_shift = kTableShift[(v / 4) % 3];
	out->join_layout = kTableLayout[(v / 4) % 3];
	out->grid = kGrids[v % 4];
	return VH_OK;
}

// Time `n` configurations of one pipeline: each is lowered and built once,
// run once untimed, then `reps` timed runs (hipEvents on `stream`); ms[i] is
// the fastest, or -1 when the configuration does not apply to the pipeline.
int vh_tune_time(const vh_pipeline_desc* desc, const vh_config* cfgs, int32_t n, int32_t reps, float* ms,
	void* stream)
{
	if (!desc || !cfgs || !ms || n <= 0 || reps <= 0)
		return fail(VH_ERR_INVALID, "vh_tune_time arguments");
	hipStream_t s = (hipStream_t)stream;
	hipEvent_t e0, e1;
	HIP_CALL(hipEventCreate(&e0));
	HIP_CALL(hipEventCreate(&e1));
	int rc = VH_OK;
	for (int i = 0; i < n && !rc; i++) {
		ms[i] = -1.f;
		vh_query* q = nullptr;
		int r = vh_query_new(&q, desc, &cfgs[i]);
		if (r == VH_ERR_UNSUPPORTED || r == VH_ERR_INVALID)
			continue;
		if (r) {
			rc = r;
			break;
		}
		float best = 3.0e38f;
		for (int k = 0; k <= reps && !rc; k++) {
			if ((rc = vh_event_record(e0, s)))
				break;
			if ((rc = vh_query_run(q, s)))
				break;
			if ((rc = vh_event_record(e1, s)))
				break;
			if (hipEventSynchronize(e1) != hipSuccess) {
				rc = fail(VH_ERR_HIP, "event synchronize");
				break;
			}
			float t = 0;
			(void)hipEventElapsedTime(&t, e0, e1);
			if (k > 0)
				best = std::min(best, t);
		}
		if (!rc) {
			// a variant whose result is an error (e.g. a full hash table) does not apply
			size_t len = 0;
			int64_t rows = 0;
			if (vh_query_result(q, stream, nullptr, 0, &len, &rows) == VH_OK)
				ms[i] = best;
		}
		vh_query_free(q);
	}
	(void)hipEventDestroy(e0);
	(void)hipEventDestroy(e1);
	return rc;
}

int vh_mlp_new(vh_mlp** out, int32_t batch_capacity, const float* init_params)
{
	if (!out || batch_capacity <= 0 || !init_params)
		return fail(VH_ERR_INVALID, "vh_mlp_new arguments");
	vh_mlp* m = new (std::nothrow) vh_mlp();
	if (!m)
		return fail(VH_ERR_NOMEM, "host allocation");
	m->cap = batch_capacity;
	const size_t pb = kParams * sizeof(float);
	const size_t ab = (size_t)batch_capacity * (kHid * 4 + kOut * 2 + 1) * sizeof(float);
	if (hipMalloc(&m->P, pb) != hipSuccess || hipMalloc(&m->G, pb) != hipSuccess ||
		hipMalloc(&m->M, pb) != hipSuccess || hipMalloc(&m->V, pb) != hipSuccess ||
		hipMalloc(&m->act, ab) != hipSuccess) {
		delete m;
		return fail(VH_ERR_NOMEM, "MLP buffers");
	}
	if (hipMemcpy(m->P, init_params, pb, hipMemcpyHostToDevice) != hipSuccess ||
		hipMemset(m->M, 0, pb) != hipSuccess || hipMemset(m->V, 0, pb) != hipSuccess ||
		hipMemset(m->G, 0, pb) != hipSuccess) {
		delete m;
		return fail(VH_ERR_HIP, "MLP initialisation");
	}
	*out = m;
	return VH_OK;
}

int vh_mlp_free(vh_mlp* m)
{
	delete m;
	return VH_OK;
}

int vh_mlp_params(vh_mlp* m, float* host_params)
{
	if (!m || !host_params)
		return fail(VH_ERR_INVALID, "vh_mlp_params arguments");
	HIP_CALL(hipMemcpy(host_params, m->P, kParams * sizeof(float), hipMemcpyDeviceToHost));
	return VH_OK;
}

// x: [B][6] device, mask: [B][48] device or NULL, prob_out: [B][48] device
int vh_mlp_predict(vh_mlp* m, const float* x, const uint8_t* mask, int32_t B, float* prob_out, void* stream)
{
	if (!m || !x || !prob_out || B <= 0 || B > m->cap)
		return fail(VH_ERR_INVALID, "vh_mlp_predict arguments");
	float* H1 = m->act;
	float* H2 = H1 + (size_t)m->cap * kHid;
	HIP_CALL(mlp_forward(m->P, x, mask, H1, H2, prob_out, B, (hipStream_t)stream));
	return VH_OK;
}

// One step on a batch: forward, softmax cross-entropy, backward into the
// gradient buffer, and (if lr > 0) an Adam update.  grads_out (host, may be
// NULL) receives the gradient; *loss the mean loss.
int vh_mlp_train_step(vh_mlp* m, const float* x, const uint8_t* mask, const int32_t* y, int32_t B, float lr,
	float* grads_out, float* loss, void* stream)
{
	if (!m || !x || !y || B <= 0 || B > m->cap)
		return fail(VH_ERR_INVALID, "vh_mlp_train_step arguments");
	hipStream_t s = (hipStream_t)stream;
	float* H1 = m->act;
	float* H2 = H1 + (size_t)m->cap * kHid;
	float* prob = H2 + (size_t)m->cap * kHid;
	float* dz = prob + (size_t)m->cap * kOut;
	float* dh2 = dz + (size_t)m->cap * kOut;
	float* dh1 = dh2 + (size_t)m->cap * kHid;
	float* lt = dh1 + (size_t)m->cap * kHid;
	HIP_CALL(mlp_forward(m->P, x, mask, H1, H2, prob, B, s));
	HIP_CALL(mlp_backward(m->P, x, H1, H2, prob, y, m->G, dz, dh2, dh1, lt, B, s));
	if (grads_out)
		HIP_CALL(hipMemcpyAsync(grads_out, m->G, kParams * sizeof(float), hipMemcpyDeviceToHost, s));
	if (loss) {
		std::vector<float> l(B);
		HIP_CALL(hipMemcpyAsync(l.data(), lt, B * sizeof(float), hipMemcpyDeviceToHost, s));
		HIP_CALL(hipStreamSynchronize(s));
		double acc = 0;
		for (float v : l)
			acc += v;
		*loss = (float)(acc / B);
	}
	if (lr > 0) {
		m->step++;
		HIP_CALL(mlp_adam_step(m->P, m->G, m->M, m->V, lr, m->step, s));
	}
	HIP_CALL(hipStreamSynchronize(s));
	return VH_OK;
}

} // extern "C"
