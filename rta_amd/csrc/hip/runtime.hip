// Runtime part of the C-ABI: errors, device memory, streams, events, and two
// utility kernels (fill, copy).  The reference reports HIP/CUDA failures by
// printing and exiting (librta/cuda.h:70-78); here every failure becomes a
// status code plus a per-thread message.
#include "common.hpp"

#include <rocprofiler-sdk-roctx/roctx.h>

#include <atomic>
#include <cstdio>
#include <mutex>
#include <string>

namespace rta_hip {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }

int hip_status(hipError_t e, const char *what) {
	g_last_error = std::string(what) + ": " + hipGetErrorString(e);
	return (int)e;
}

int invalid(const char *msg) {
	g_last_error = msg;
	return RTA_E_INVALID;
}

// One-shot fault injection (rta_hip_inject_fault): the armed entry point's
// next call returns the chosen status before doing any work.
static std::mutex g_fault_mutex;
static std::atomic<int> g_fault_armed{0};
static std::string g_fault_entry;
static int g_fault_status = 0;

bool take_fault(const char *entry, int *status) {
	if (g_fault_armed.load(std::memory_order_relaxed) == 0) return false;
	std::lock_guard<std::mutex> lock(g_fault_mutex);
	if (g_fault_armed.load(std::memory_order_relaxed) == 0 || g_fault_entry != entry) return false;
	g_fault_armed.store(0, std::memory_order_relaxed);
	*status = g_fault_status;
	g_last_error = std::string(entry) + ": injected fault";
	return true;
}

__global__ void fill_u32_kernel(uint32_t *dst, uint32_t v, size_t n) {
	size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	size_t stride = (size_t)gridDim.x * blockDim.x;
	for (; i < n; i += stride) dst[i] = v;
}

__global__ void add_u32_kernel(uint32_t *word, uint32_t delta) {
	if (threadIdx.x == 0) *word += delta;
}

__global__ void copy16_kernel(uint4 *__restrict__ dst, const uint4 *__restrict__ src, size_t n) {
	size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	size_t stride = (size_t)gridDim.x * blockDim.x;
	for (; i < n; i += stride) dst[i] = src[i];
}

}  // namespace rta_hip

using namespace rta_hip;

extern "C" {

const char *rta_hip_last_error(void) { return g_last_error.c_str(); }

int rta_hip_range_push(const char *name) {
	if (!name) return invalid("rta_hip_range_push: name is NULL");
	roctxRangePushA(name);
	return RTA_OK;
}

int rta_hip_range_pop(void) {
	roctxRangePop();
	return RTA_OK;
}

int rta_hip_inject_fault(const char *entry, int status) {
	if (!entry || !*entry || status == RTA_OK)
		return invalid("rta_hip_inject_fault: an entry point name and a non-zero status");
	std::lock_guard<std::mutex> lock(g_fault_mutex);
	g_fault_entry = entry;
	g_fault_status = status;
	g_fault_armed.store(1, std::memory_order_relaxed);
	return RTA_OK;
}

int rta_hip_device_count(int *count) {
	if (!count) return invalid("rta_hip_device_count: count is NULL");
	RTA_HIP_TRY(hipGetDeviceCount(count));
	return RTA_OK;
}

int rta_hip_set_device(int device) {
	RTA_HIP_TRY(hipSetDevice(device));
	return RTA_OK;
}

int rta_hip_device_name(char *buf, int buflen) {
	if (!buf || buflen <= 0) return invalid("rta_hip_device_name: bad buffer");
	int dev = 0;
	RTA_HIP_TRY(hipGetDevice(&dev));
	hipDeviceProp_t prop;
	RTA_HIP_TRY(hipGetDeviceProperties(&prop, dev));
	snprintf(buf, (size_t)buflen, "%s (%s, %d CUs)", prop.name, prop.gcnArchName,
	         prop.multiProcessorCount);
	return RTA_OK;
}

int rta_hip_malloc(void **ptr, size_t bytes) {
	RTA_FAULT_POINT("rta_hip_malloc");
	if (!ptr) return invalid("rta_hip_malloc: ptr is NULL");
	*ptr = nullptr;
	if (bytes == 0) return RTA_OK;
	RTA_HIP_TRY(hipMalloc(ptr, bytes));
	return RTA_OK;
}

int rta_hip_free(void *ptr) {
	if (!ptr) return RTA_OK;
	RTA_HIP_TRY(hipFree(ptr));
	return RTA_OK;
}

int rta_hip_memcpy_htod(void *dst, const void *src, size_t bytes) {
	if (bytes == 0) return RTA_OK;
	if (!dst || !src) return invalid("rta_hip_memcpy_htod: NULL pointer");
	RTA_HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
	return RTA_OK;
}

int rta_hip_memcpy_dtoh(void *dst, const void *src, size_t bytes) {
	if (bytes == 0) return RTA_OK;
	if (!dst || !src) return invalid("rta_hip_memcpy_dtoh: NULL pointer");
	RTA_HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
	return RTA_OK;
}

int rta_hip_memcpy_dtod(void *dst, const void *src, size_t bytes, rta_stream s) {
	if (bytes == 0) return RTA_OK;
	if (!dst || !src) return invalid("rta_hip_memcpy_dtod: NULL pointer");
	RTA_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, as_stream(s)));
	return RTA_OK;
}

int rta_hip_fill_u32(uint32_t *dst, uint32_t value, size_t count, rta_stream s) {
	if (count == 0) return RTA_OK;
	if (!dst) return invalid("rta_hip_fill_u32: dst is NULL");
	unsigned blocks = grid_for((long long)count, 256);
	if (blocks > 4096) blocks = 4096;
	hipLaunchKernelGGL(fill_u32_kernel, dim3(blocks), dim3(256), 0, as_stream(s), dst, value, count);
	RTA_LAUNCH_CHECK("fill_u32_kernel");
	return RTA_OK;
}

int rta_hip_add_u32(uint32_t *word, uint32_t delta, rta_stream s) {
	if (!word) return invalid("rta_hip_add_u32: word is NULL");
	hipLaunchKernelGGL(add_u32_kernel, dim3(1), dim3(64), 0, as_stream(s), word, delta);
	RTA_LAUNCH_CHECK("add_u32_kernel");
	return RTA_OK;
}

int rta_hip_copy_kernel(void *dst, const void *src, size_t bytes, rta_stream s) {
	if (bytes == 0) return RTA_OK;
	if (!dst || !src || (bytes & 15)) return invalid("rta_hip_copy_kernel: bad arguments");
	size_t n = bytes / 16;
	unsigned blocks = grid_for((long long)n, 256);
	if (blocks > 8192) blocks = 8192;
	hipLaunchKernelGGL(copy16_kernel, dim3(blocks), dim3(256), 0, as_stream(s),
	                   (uint4 *)dst, (const uint4 *)src, n);
	RTA_LAUNCH_CHECK("copy16_kernel");
	return RTA_OK;
}

int rta_hip_stream_create(rta_stream *s) {
	if (!s) return invalid("rta_hip_stream_create: NULL");
	hipStream_t st;
	RTA_HIP_TRY(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
	*s = (rta_stream)st;
	return RTA_OK;
}

int rta_hip_stream_destroy(rta_stream s) {
	if (!s) return RTA_OK;
	RTA_HIP_TRY(hipStreamDestroy(as_stream(s)));
	return RTA_OK;
}

int rta_hip_stream_synchronize(rta_stream s) {
	RTA_HIP_TRY(hipStreamSynchronize(as_stream(s)));
	return RTA_OK;
}

int rta_hip_device_synchronize(void) {
	RTA_HIP_TRY(hipDeviceSynchronize());
	return RTA_OK;
}

int rta_hip_event_create(rta_event *e) {
	if (!e) return invalid("rta_hip_event_create: NULL");
	hipEvent_t ev;
	RTA_HIP_TRY(hipEventCreate(&ev));
	*e = (rta_event)ev;
	return RTA_OK;
}

int rta_hip_event_destroy(rta_event e) {
	if (!e) return RTA_OK;
	RTA_HIP_TRY(hipEventDestroy((hipEvent_t)e));
	return RTA_OK;
}

int rta_hip_event_record(rta_event e, rta_stream s) {
	if (!e) return invalid("rta_hip_event_record: NULL event");
	RTA_HIP_TRY(hipEventRecord((hipEvent_t)e, as_stream(s)));
	return RTA_OK;
}

int rta_hip_stream_wait_event(rta_stream s, rta_event e) {
	if (!e) return invalid("rta_hip_stream_wait_event: NULL event");
	RTA_HIP_TRY(hipStreamWaitEvent(as_stream(s), (hipEvent_t)e, 0));
	return RTA_OK;
}

int rta_hip_event_synchronize(rta_event e) {
	if (!e) return invalid("rta_hip_event_synchronize: NULL event");
	RTA_HIP_TRY(hipEventSynchronize((hipEvent_t)e));
	return RTA_OK;
}

int rta_hip_event_elapsed_ms(float *ms, rta_event start, rta_event stop) {
	if (!ms || !start || !stop) return invalid("rta_hip_event_elapsed_ms: NULL");
	RTA_HIP_TRY(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop));
	return RTA_OK;
}

// A captured graph and its executable form.
struct rta_graph_rec {
	hipGraph_t graph;
	hipGraphExec_t exec;
};

int rta_hip_graph_begin(rta_stream s) {
	if (!s) return invalid("rta_hip_graph_begin: capture needs a created stream, not NULL");
	RTA_HIP_TRY(hipStreamBeginCapture(as_stream(s), hipStreamCaptureModeThreadLocal));
	return RTA_OK;
}

int rta_hip_graph_end(rta_stream s, rta_graph *g) {
	if (!s || !g) return invalid("rta_hip_graph_end: NULL argument");
	hipGraph_t graph = nullptr;
	RTA_HIP_TRY(hipStreamEndCapture(as_stream(s), &graph));
	hipGraphExec_t exec = nullptr;
	hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
	if (e != hipSuccess) {
		(void)hipGraphDestroy(graph);
		return hip_status(e, "hipGraphInstantiate");
	}
	*g = (rta_graph) new rta_graph_rec{graph, exec};
	return RTA_OK;
}

int rta_hip_graph_launch(rta_graph g, rta_stream s) {
	if (!g) return invalid("rta_hip_graph_launch: NULL graph");
	RTA_HIP_TRY(hipGraphLaunch(((rta_graph_rec *)g)->exec, as_stream(s)));
	return RTA_OK;
}

int rta_hip_graph_destroy(rta_graph g) {
	if (!g) return RTA_OK;
	rta_graph_rec *r = (rta_graph_rec *)g;
	hipError_t e1 = hipGraphExecDestroy(r->exec);
	hipError_t e2 = hipGraphDestroy(r->graph);
	delete r;
	if (e1 != hipSuccess) return hip_status(e1, "hipGraphExecDestroy");
	if (e2 != hipSuccess) return hip_status(e2, "hipGraphDestroy");
	return RTA_OK;
}

}  // extern "C"
