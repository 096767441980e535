// The LBVH build for callers that cannot allocate device memory (code under
// graph capture, tight serving loops, harnesses that forbid hipMalloc):
// rta_lbvh_build with its workspace in a file-scope device array instead of
// a caller buffer.  No counterpart in the reference.
//
// The array is `scratch` below.  In host code that name is a host-side
// stand-in, not a device address: the device address is taken once per
// device with hipGetSymbolAddress (each GPU holds its own copy of the array
// at its own address) and only that pointer is handed to kernels.  Every
// call on one GPU shares that GPU's array, so calls that overlap on one GPU
// (different streams, or host threads) race; calls in one stream do not.
#include "common.hpp"

#include <mutex>

namespace rta_hip {

// Sized for RTA_LBVH_NOALLOC_MAX_TRIS triangles: rta_lbvh_workspace_bytes
// gives 44,199,424 bytes at 262,144 triangles (168.6 bytes per triangle);
// the first call checks it against the array.
constexpr size_t kNoAllocScratchBytes = (size_t)46 << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mutex;
static void *g_scratch_ptr[kMaxDevices];

// The current device's address of `scratch`, looked up on the first call
// on that device (with that device current).
static int scratch_ptr(void **out) {
	int dev = -1;
	RTA_HIP_TRY(hipGetDevice(&dev));
	if (dev < 0 || dev >= kMaxDevices)
		return invalid("rta_lbvh_build_noalloc: device index beyond 63");
	std::lock_guard<std::mutex> lock(g_scratch_mutex);
	if (!g_scratch_ptr[dev]) {
		size_t need = 0;
		const int rc = rta_lbvh_workspace_bytes(RTA_LBVH_NOALLOC_MAX_TRIS, &need);
		if (rc != RTA_OK) return rc;
		if (need > kNoAllocScratchBytes)
			return invalid("rta_lbvh_build_noalloc: static workspace smaller than the limit needs");
		void *p = nullptr;
		RTA_HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
		if (!p) return invalid("rta_lbvh_build_noalloc: no device address for the workspace");
		g_scratch_ptr[dev] = p;
	}
	*out = g_scratch_ptr[dev];
	return RTA_OK;
}

}  // namespace rta_hip

using namespace rta_hip;

extern "C" {

int rta_lbvh_build_noalloc(rta_triangle *tris, int n, rta_node *nodes, uint32_t *indices,
                           uint32_t *status, rta_stream s) {
	if (n <= 0) return invalid("rta_lbvh_build_noalloc: n must be > 0");
	if (!tris || !nodes || !indices) return invalid("rta_lbvh_build_noalloc: NULL pointer");
	if (n > RTA_LBVH_NOALLOC_MAX_TRIS) {
		set_error("rta_lbvh_build_noalloc: n exceeds RTA_LBVH_NOALLOC_MAX_TRIS (262144); "
		          "use rta_lbvh_build with a workspace");
		return RTA_E_WORKSPACE;
	}
	void *ws = nullptr;
	const int rc = scratch_ptr(&ws);
	if (rc != RTA_OK) return rc;
	// every workspace word a stage reads is written earlier in the same
	// build (the bottom-up visit counters and frontier status are cleared
	// by the hierarchy stage), so nothing left by a previous call matters
	return rta_lbvh_build(tris, n, nodes, indices, ws, kNoAllocScratchBytes, status, nullptr, s);
}

}  // extern "C"
