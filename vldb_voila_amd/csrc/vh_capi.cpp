// extern "C" entry points of libvoila_hip.so (declared in include/voila_hip.h).
#include <hip/hip_runtime.h>

#include <cstring>
#include <new>

#include "../../include/voila_hip.h"
#include "vh_engine.hpp"
#include "vh_jit.hpp"

using namespace vh;

struct vh_query {
	Query q;
};

#define HIP_CALL(x)                                                                  \
	do {                                                                             \
		hipError_t _e = (x);                                                         \
		if (_e != hipSuccess)                                                        \
			return fail(VH_ERR_HIP, "%s: %s", #x, hipGetErrorString(_e));           \
	} while (0)

extern "C" {

const char* vh_last_error(void) { return last_error(); }

int vh_init(int device)
{
	int n = 0;
	if (hipGetDeviceCount(&n) != hipSuccess || n == 0)
		return fail(VH_ERR_NODEVICE, "no HIP device");
	if (device < 0 || device >= n)
		return fail(VH_ERR_INVALID, "device %d of %d", device, n);
	HIP_CALL(hipSetDevice(device));
	return VH_OK;
}

int vh_malloc(void** ptr, size_t bytes)
{
	if (!ptr)
		return fail(VH_ERR_INVALID, "null ptr");
	hipError_t e = hipMalloc(ptr, bytes ? bytes : 1);
	if (e != hipSuccess)
		return fail(VH_ERR_NOMEM, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
	return VH_OK;
}

int vh_free(void* ptr)
{
	HIP_CALL(hipFree(ptr));
	return VH_OK;
}

int vh_memcpy_h2d(void* dst, const void* src, size_t bytes)
{
	HIP_CALL(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
	return VH_OK;
}

int vh_memcpy_d2h(void* dst, const void* src, size_t bytes)
{
	HIP_CALL(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
	return VH_OK;
}

int vh_synchronize(void)
{
	HIP_CALL(hipDeviceSynchronize());
	return VH_OK;
}

int vh_stream_create(void** stream)
{
	HIP_CALL(hipStreamCreate((hipStream_t*)stream));
	return VH_OK;
}

int vh_stream_destroy(void* stream)
{
	HIP_CALL(hipStreamDestroy((hipStream_t)stream));
	return VH_OK;
}

int vh_stream_synchronize(void* stream)
{
	HIP_CALL(hipStreamSynchronize((hipStream_t)stream));
	return VH_OK;
}

int vh_event_create(void** event)
{
	HIP_CALL(hipEventCreate((hipEvent_t*)event));
	return VH_OK;
}

int vh_event_destroy(void* event)
{
	HIP_CALL(hipEventDestroy((hipEvent_t)event));
	return VH_OK;
}

int vh_event_record(void* event, void* stream)
{
	HIP_CALL(hipEventRecord((hipEvent_t)event, (hipStream_t)stream));
	return VH_OK;
}

int vh_event_elapsed_ms(float* ms, void* start, void* stop)
{
	HIP_CALL(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop));
	return VH_OK;
}

int vh_type_from_minmax(double min, double max) { return type_from_minmax(min, max); }

int vh_query_new(vh_query** out, const vh_pipeline_desc* desc, const vh_config* cfg)
{
	if (!out || !desc)
		return fail(VH_ERR_INVALID, "null argument");
	*out = nullptr;
	vh_query* h = new (std::nothrow) vh_query();
	if (!h)
		return fail(VH_ERR_NOMEM, "host allocation");
	memcpy(&h->q.desc, desc, sizeof(*desc));
	if (cfg) {
		h->q.cfg = *cfg;
	} else {
		memset(&h->q.cfg, 0, sizeof(h->q.cfg));
		h->q.cfg.group_mode = VH_GROUP_AUTO;
	}
	int rc = lower(&h->q.desc, &h->q.cfg, h->q.L);
	if (!rc)
		rc = query_create(&h->q);
	if (rc) {
		delete h;
		return rc;
	}
	*out = h;
	return VH_OK;
}

int vh_query_run(vh_query* q, void* stream)
{
	if (!q)
		return fail(VH_ERR_INVALID, "null query");
	return query_run(&q->q, (hipStream_t)stream, -1);
}

int vh_query_run_phase(vh_query* q, void* stream, int phase)
{
	if (!q)
		return fail(VH_ERR_INVALID, "null query");
	return query_run(&q->q, (hipStream_t)stream, phase);
}

int vh_query_result(vh_query* q, void* stream, char* buf, size_t cap, size_t* len, int64_t* nrows)
{
	if (!q || !len)
		return fail(VH_ERR_INVALID, "null argument");
	int64_t rows = 0;
	int rc = query_result(&q->q, (hipStream_t)stream, q->q.result, rows);
	if (rc)
		return rc;
	*len = q->q.result.size();
	if (nrows)
		*nrows = rows;
	if (buf) {
		if (cap <= q->q.result.size())
			return fail(VH_ERR_CAPACITY, "result needs %zu bytes", q->q.result.size() + 1);
		memcpy(buf, q->q.result.data(), q->q.result.size());
		buf[q->q.result.size()] = 0;
	}
	return VH_OK;
}

int vh_query_reset(vh_query* q)
{
	if (!q)
		return fail(VH_ERR_INVALID, "null query");
	q->q.ran = false;
	q->q.result.clear();
	return VH_OK;
}

int vh_query_free(vh_query* q)
{
	delete q;
	return VH_OK;
}

int vh_query_info(const vh_query* q, int32_t* group_mode, int32_t* grid, int64_t* groups)
{
	if (!q)
		return fail(VH_ERR_INVALID, "null query");
	if (group_mode)
		*group_mode = q->q.L.plan.mode;
	if (grid)
		*grid = q->q.L.grid;
	if (groups)
		*groups = (int64_t)q->q.L.groups;
	return VH_OK;
}

int vh_query_join(const vh_query* q, int32_t j, int32_t* entry_bytes, int64_t* entries)
{
	if (!q || j < 0 || j >= q->q.L.plan.njoins)
		return fail(VH_ERR_INVALID, "vh_query_join arguments");
	const BuildPlan& B = q->q.L.builds[j];
	const bool direct = B.dir_w != 0;
	if (entry_bytes)
		*entry_bytes = direct ? B.dir_w : 8;
	if (entries)
		*entries = direct ? (int64_t)(B.kmax - B.kmin + 1) : (int64_t)q->q.L.capacity[j];
	return VH_OK;
}

int vh_generate_source(const vh_pipeline_desc* desc, const vh_config* cfg, char* buf, size_t cap, size_t* len)
{
	if (!desc || !len)
		return fail(VH_ERR_INVALID, "vh_generate_source arguments");
	Lowered L;
	int rc = lower(desc, cfg, L);
	if (rc)
		return rc;
	const int kernel = cfg ? cfg->kernel : VH_KERNEL_AUTO;
	if (kernel != VH_KERNEL_GENERIC)
		choose_probe_shape(L, desc, cfg);
	std::string src = jit_source(L.plan, jit_eager(kernel), L.plan.radix ? L.radix_unroll : jit_unroll(cfg),
		jit_pipeline(kernel), L.block);
	if (L.plan.radix)
		src += jit_join_source(L.plan, jit_unroll(cfg), L.radix_block);
	*len = src.size();
	if (buf) {
		if (cap <= src.size())
			return fail(VH_ERR_CAPACITY, "source needs %zu bytes", src.size() + 1);
		memcpy(buf, src.data(), src.size());
		buf[src.size()] = 0;
	}
	return VH_OK;
}

int vh_query_probe_events(vh_query* q, void* start, void* stop)
{
	if (!q)
		return fail(VH_ERR_INVALID, "null query");
	if (!q->q.jit_fn && (start || stop))
		return fail(VH_ERR_UNSUPPORTED, "kernel timestamps are recorded for the generated probe kernel only");
	q->q.probe_start = (hipEvent_t)start;
	q->q.probe_stop = (hipEvent_t)stop;
	return VH_OK;
}

int vh_query_kernel(const vh_query* q, int32_t* kernel, char* buf, size_t cap, size_t* len)
{
	if (!q)
		return fail(VH_ERR_INVALID, "null query");
	// a radix-partitioned probe runs two generated kernels: both sources
	const std::string src = q->q.jit_src + q->q.jit_join_src;
	if (kernel)
		*kernel = q->q.jit_fn ? VH_KERNEL_JIT : VH_KERNEL_GENERIC;
	if (len)
		*len = src.size();
	if (buf) {
		if (cap <= src.size())
			return fail(VH_ERR_CAPACITY, "source needs %zu bytes", src.size() + 1);
		memcpy(buf, src.data(), src.size());
		buf[src.size()] = 0;
	}
	return VH_OK;
}

} // extern "C"
