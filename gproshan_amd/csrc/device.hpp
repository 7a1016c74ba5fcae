// Device-side declarations shared by the HIP translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdlib>

#include "gp_internal.hpp"

namespace gp {

#define GP_HIP(call)                                                                  \
	do {                                                                              \
		hipError_t e_ = (call);                                                       \
		if (e_ != hipSuccess)                                                         \
			return fail(GP_ERR_HIP, "%s failed: %s", #call, hipGetErrorString(e_)); \
	} while (0)

static_assert(sizeof(Geo32) == sizeof(float2), "Geo32 must match float2");

struct MasterLayout;  // ptp_gpu.hip

struct DeviceMesh {
	uint32_t nv = 0, nfan = 0;
	uint32_t *fan_off = nullptr;
	uint32_t *fan = nullptr;   // fan lists (vertex ids)
	double *pos = nullptr;     // 3*nv
	uint8_t *closed = nullptr; // 1 when the fan of v is closed (link = first m entries)
	float h = 1.f, inv_h = 1.f;
	MasterLayout *master[2] = { nullptr, nullptr };   // vertex-major slot tables (FP32, FP64), built on demand
};

// Uploads the fan tables and FP64 positions of m once (m.dev caches them).
int upload_mesh(Che &m);

// Replayed launch sequences (sweeps, BFS levels, patch launches, the batch)
// run from hipGraphs unless the environment sets
// GPROSHAN_AMD_GRAPHS=0: then the same kernels are launched one by one on the
// stream, with the same results (a profiler that cannot follow graph replays
// traces these).
inline bool graphs_enabled()
{
	static const bool on = [] {
		const char *e = std::getenv("GPROSHAN_AMD_GRAPHS");
		return !(e && e[0] == '0');
	}();
	return on;
}

// A chain of kernel launches built as a graph with the explicit node API.
// Stream capture is not used: while one thread captures, HIP fails another
// thread's legacy-stream calls (hipMemcpy and the like) and that error also
// spoils the capture, so plans built from several threads would break each
// other.  A node copies its arguments when it is added.
struct GraphChain {
	hipGraph_t g = nullptr;
	hipGraphNode_t last = nullptr;
	hipError_t err = hipSuccess;
	GraphChain() { err = hipGraphCreate(&g, 0); }
	~GraphChain()
	{
		if (g) (void) hipGraphDestroy(g);
	}
	GraphChain(const GraphChain &) = delete;
	GraphChain &operator=(const GraphChain &) = delete;
	void add(const void *fn, dim3 grid, dim3 block, void **args)
	{
		if (err != hipSuccess) return;
		hipKernelNodeParams kp{};
		kp.func = const_cast<void *>(fn);
		kp.gridDim = grid;
		kp.blockDim = block;
		kp.sharedMemBytes = 0;
		kp.kernelParams = args;
		kp.extra = nullptr;
		hipGraphNode_t n = nullptr;
		err = hipGraphAddKernelNode(&n, g, last ? &last : nullptr, last ? 1 : 0, &kp);
		if (err == hipSuccess) last = n;
	}
	hipError_t instantiate(hipGraphExec_t *out)
	{
		if (err != hipSuccess) return err;
		return hipGraphInstantiate(out, g, nullptr, nullptr, 0);
	}
};

// Copies a state record of n 32-bit words from device memory into pinned
// host memory (hipHostMalloc, mapped into the device's address space); the
// host reads it after an event recorded behind this launch.  The replay loops
// take their per-replay snapshots this way instead of with hipMemcpyAsync.
static __global__ void k_snapshot_words(const uint32_t *src, uint32_t *dst, uint32_t n)
{
	for (uint32_t q = threadIdx.x; q < n; q += blockDim.x) dst[q] = src[q];
}

template <typename S>
inline void snapshot(const S *dev, S *host, hipStream_t stream)
{
	static_assert(sizeof(S) % 4 == 0, "state records are whole words");
	k_snapshot_words<<<1, 64, 0, stream>>>((const uint32_t *) dev, (uint32_t *) host, (uint32_t) (sizeof(S) / 4));
}


}  // namespace gp
