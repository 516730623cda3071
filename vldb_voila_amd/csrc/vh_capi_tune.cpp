// extern "C" entry points of the autotuner (exhaustive variant timing) and
// the learned selector (MLP training and inference on the GPU).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <string>
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

// Variant index -> knobs: 4 group-by strategies x 3 join-table layouts and
// partition fan-outs x 4 launch shapes = 48, the analog of the explorer's
// blend space (explorer_helper.cpp:6-17: computation type x concurrent FSMs
// x prefetch).
static const int kModes[4] = {VH_GROUP_DENSE, VH_GROUP_LDSHASH, VH_GROUP_PART, VH_GROUP_HASH};
// join tables: {direct when the keys are dense, built partitioned by key
// range (else a hash table, partitioned above 4 MiB)}, {hash table built in
// 2^13-slot partitions, fan-out = slots / 8192}, {one hash table, global
// CAS build, no partitions}
static const int kLayout[3] = {VH_LAYOUT_AUTO, VH_LAYOUT_HASH, VH_LAYOUT_HASH};
static const int kPartShift[3] = {0, 13, -1};
// launch shapes: {threads per workgroup, rows per lane (tile = threads x
// rows), workgroups (0 = the lowering's: two 256-thread ones per CU),
// batched work item 16384 << tile_shift rows, 8192 << tile_shift for a
// partitioned group-by}.  Waves per CU: 8, 16, 8, 16.
static const int kShape[4][4] = {{256, 8, 0, 0}, {256, 2, 1024, -1}, {512, 4, 0, 0}, {1024, 4, 0, 1}};

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
	out->join_layout = kLayout[(v / 4) % 3];
	out->part_shift = kPartShift[(v / 4) % 3];
	out->block = kShape[v % 4][0];
	out->unroll = kShape[v % 4][1];
	out->grid = kShape[v % 4][2];
	out->tile_shift = kShape[v % 4][3];
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

// ---- tuning cache ----------------------------------------------------------

// What the best variant of a pipeline depends on: its column types, the
// numbers of predicates, joins, keys and sums, each join's build side
// (rows, key range, payload columns, filtered or not) and the probe rows and
// group domain, sizes bucketed by powers of two.  Pipelines with the same
// signature share a tuning result.
static std::string tune_signature(const vh_pipeline_desc* d)
{
	auto bucket = [](double x) {
		int b = 0;
		while (b < 63 && (double)(1ull << b) < x)
			b++;
		return b;
	};
	std::string s = "p" + std::to_string(bucket((double)d->probe.rows)) + "c";
	for (int c = 0; c < d->probe.ncols; c++)
		s += std::to_string(d->probe.cols[c].type);
	s += "f" + std::to_string(d->probe.npreds);
	for (int j = 0; j < d->njoins; j++) {
		const vh_join_desc& J = d->joins[j];
		const vh_column& k = J.build.cols[J.build_key];
		s += "|j" + std::to_string(bucket((double)J.build.rows)) + "r" +
			std::to_string(bucket((double)k.max - (double)k.min + 1)) + "k" + std::to_string(std::max(J.nkeys, 1)) +
			"p" + std::to_string(J.npayload) + "v" + std::to_string(J.variant) + "f" + std::to_string(J.build.npreds);
	}
	double domain = 1;
	Typing ty;
	if (type_pipeline(d, ty) == VH_OK)
		for (int k = 0; k < d->nkeys; k++)
			domain *= ty.vcols[d->keys[k]].dmax - ty.vcols[d->keys[k]].dmin + 1;
	s += "|g" + std::to_string(d->nkeys) + "d" + std::to_string(bucket(domain)) + "a" + std::to_string(d->naggrs);
	return s;
}

static std::mutex g_tune_mu;
static std::map<std::string, int> g_tune_cache;

int vh_tune_signature(const vh_pipeline_desc* desc, char* buf, size_t cap, size_t* len)
{
	if (!desc || !len)
		return fail(VH_ERR_INVALID, "vh_tune_signature arguments");
	const std::string s = tune_signature(desc);
	*len = s.size();
	if (buf) {
		if (cap <= s.size())
			return fail(VH_ERR_CAPACITY, "signature needs %zu bytes", s.size() + 1);
		memcpy(buf, s.data(), s.size() + 1);
	}
	return VH_OK;
}

int vh_tune_cache_put(const char* signature, int32_t variant)
{
	if (!signature || variant < 0)
		return fail(VH_ERR_INVALID, "vh_tune_cache_put arguments");
	std::lock_guard<std::mutex> lock(g_tune_mu);
	g_tune_cache[signature] = variant;
	return VH_OK;
}

int vh_tune_cache_get(const char* signature, int32_t* variant)
{
	if (!signature || !variant)
		return fail(VH_ERR_INVALID, "vh_tune_cache_get arguments");
	std::lock_guard<std::mutex> lock(g_tune_mu);
	auto it = g_tune_cache.find(signature);
	*variant = it == g_tune_cache.end() ? -1 : it->second;
	return VH_OK;
}

int vh_tune_cache_clear(void)
{
	std::lock_guard<std::mutex> lock(g_tune_mu);
	g_tune_cache.clear();
	return VH_OK;
}

int vh_tune_best(const vh_pipeline_desc* desc, const vh_config* cfgs, int32_t n, int32_t reps, int32_t* best,
	int32_t* cached, void* stream)
{
	if (!desc || !cfgs || !best || n <= 0 || reps <= 0)
		return fail(VH_ERR_INVALID, "vh_tune_best arguments");
	const std::string sig = tune_signature(desc);
	{
		std::lock_guard<std::mutex> lock(g_tune_mu);
		auto it = g_tune_cache.find(sig);
		if (it != g_tune_cache.end() && it->second < n) {
			*best = it->second;
			if (cached)
				*cached = 1;
			return VH_OK;
		}
	}
	std::vector<float> ms((size_t)n);
	int rc = vh_tune_time(desc, cfgs, n, reps, ms.data(), stream);
	if (rc)
		return rc;
	int b = -1;
	for (int i = 0; i < n; i++)
		if (ms[i] >= 0 && (b < 0 || ms[i] < ms[b]))
			b = i;
	if (b < 0)
		return fail(VH_ERR_UNSUPPORTED, "no variant applies to the pipeline");
	{
		std::lock_guard<std::mutex> lock(g_tune_mu);
		g_tune_cache[sig] = b;
	}
	*best = b;
	if (cached)
		*cached = 0;
	return VH_OK;
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
