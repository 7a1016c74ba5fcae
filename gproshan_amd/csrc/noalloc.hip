// The allocation-free variant of parallel_toplesets_propagation_gpu
// (src/cuda/geodesics_ptp.cu:20-85,174-232): the reference window schedule in
// FP64 over sorted[], for callers that cannot allocate device memory (code
// under HIP graph capture, serving loops).  Inputs and the output are the
// caller's device buffers; the schedule state and the two distance buffers
// live in one file-scope device array per GPU (`scratch`), whose address is
// looked up once per device; every kernel is enqueued on the caller's stream
// and nothing is allocated, freed or synchronised.
//
// The sweeps relax vertex sorted[p] for window position p through the mesh's
// vertex-major FP64 slot tables (slot k of v at fan_off[v] + k; the same fan
// order and update_step operations as every FP64 solve, hence the same bits
// as gp_parallel_toplesets_propagation_gpu).  Those tables are built on the
// device by the first FP64 solve with device_setup on the mesh; without them
// the call is rejected.  The schedule runs its full 2 * |limits| + 1 launches
// (the reference's cap plus the launch that records the last outcome):
// launches after the one that finishes see the done state and return at once,
// so the host never reads the state.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>

#include "device.hpp"
#include "sweep_kernel.hpp"

namespace gp {

// Schedule state, then the two FP64 distance buffers of at most
// NOALLOC_MAX_VERTICES vertices each.
constexpr uint32_t NOALLOC_MAX_VERTICES = 1u << 20;
constexpr size_t NOALLOC_STATE_BYTES = (sizeof(PState) + 255) & ~size_t(255);
constexpr size_t NOALLOC_SCRATCH_BYTES = NOALLOC_STATE_BYTES + 2 * sizeof(double) * NOALLOC_MAX_VERTICES;
static_assert(GP_NOALLOC_MAX_VERTICES == NOALLOC_MAX_VERTICES, "header and library agree on the limit");

static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int MAX_DEVICES = 64;

struct DeviceScratch {
	unsigned char *ptr = nullptr;   // the device address of `scratch` on this device
	uint32_t grid = 0;              // sweep workgroups that fit on the device at once
};

std::mutex g_mu;
DeviceScratch g_dev[MAX_DEVICES];

// The scratch address and launch width of the current device, looked up on
// the first call made with that device current.
int device_scratch(DeviceScratch &out)
{
	int dev = 0;
	GP_HIP(hipGetDevice(&dev));
	if (dev < 0 || dev >= MAX_DEVICES) return fail(GP_ERR_CAPACITY, "noalloc: device index %d", dev);
	std::lock_guard<std::mutex> lock(g_mu);
	DeviceScratch &d = g_dev[dev];
	if (!d.ptr) {
		void *p = nullptr;
		const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
		if (e != hipSuccess || !p)
			return fail(GP_ERR_HIP, "noalloc: hipGetSymbolAddress failed: %s", hipGetErrorString(e));
		int ncu = 0, occ = 0;
		GP_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
		GP_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, (const void *) k_sweep<true, false, MAXS, false, true>,
		                                                    (int) SWEEP_BLOCK, 0));
		d.grid = (uint32_t) ncu * (uint32_t) std::max(1, occ);
		d.ptr = (unsigned char *) p;
	}
	out = d;
	return GP_OK;
}

// The buffer the last sweep wrote (the GPU drivers' result,
// src/cuda/geodesics_ptp.cu:228-231) into the caller's output.
__global__ void k_noalloc_result(const double *d0, const double *d1, const PState *st, uint32_t n, double *out)
{
	const double *src = st->final_d ? d1 : d0;
	for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) out[v] = src[v];
}

// Sources at distance 0 in both buffers (ids at or past n are ignored: the
// device inputs are not checked on the host).
__global__ void k_noalloc_sources(double *d0, double *d1, const uint32_t *src, uint32_t ns, uint32_t n)
{
	const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
	if (t < ns && src[t] < n) { d0[src[t]] = 0.0; d1[src[t]] = 0.0; }
}

}  // namespace

int ptp_gpu_noalloc(Che &mesh, const uint32_t *d_src, uint32_t ns, const uint32_t *d_limits, uint32_t nl,
                    const uint32_t *d_sorted, double *d_dist, void *stream_handle)
{
	const hipStream_t stream = (hipStream_t) stream_handle;
	if (!d_src || !d_limits || !d_sorted || !d_dist) return fail(GP_ERR_INVALID, "noalloc: null argument");
	if (ns == 0 || nl < 2) return fail(GP_ERR_INVALID, "noalloc: no sources / empty rings");
	const uint32_t n = mesh.nv;
	if (n > NOALLOC_MAX_VERTICES)
		return fail(GP_ERR_CAPACITY, "noalloc: %u vertices, at most %u", n, NOALLOC_MAX_VERTICES);
	const MasterLayout *ml = mesh.dev ? mesh.dev->master[1] : nullptr;
	if (!ml)
		return fail(GP_ERR_INVALID, "noalloc: the mesh's FP64 device tables are not built (run one FP64 solve with "
		                            "device_setup on this mesh first)");
	DeviceScratch ds;
	int rc = device_scratch(ds);
	if (rc) return rc;
	PState *st = (PState *) ds.ptr;
	double *d0 = (double *) (ds.ptr + NOALLOC_STATE_BYTES);
	double *d1 = d0 + NOALLOC_MAX_VERTICES;
	const VertexTables vt = ml->view(mesh.dev);
	KArgs a{};
	a.limits = d_limits;
	a.coff = vt.coff;
	a.cid = vt.cid;
	a.cx64 = vt.cx64;
	a.order = d_sorted;
	a.dist0 = d0;
	a.dist1 = d1;
	a.st = st;
	a.tol = GP_PTP_TOL;
	a.nl = nl;
	a.max_iter = nl << 1;                 // src/geodesics_ptp.cpp:66-67
	a.npos = n;
	a.k2 = 1;
	a.closed = vt.closed;
	const uint32_t ig = std::max<uint32_t>(1, std::min<uint32_t>(1024, (n + 255) / 256));
	k_init<double><<<ig, 256, 0, stream>>>(d0, d1, n, st, nl < 3 ? 1u : 0u);
	k_noalloc_sources<<<(ns + 255) / 256, 256, 0, stream>>>(d0, d1, d_src, ns, n);
	for (uint32_t k = 0; k <= a.max_iter; k++)
		k_sweep<true, false, MAXS, false, true><<<ds.grid, SWEEP_BLOCK, 0, stream>>>(a.st, k & 3u, a);
	k_noalloc_result<<<ig, 256, 0, stream>>>(d0, d1, st, n, d_dist);
	GP_HIP(hipGetLastError());
	return GP_OK;
}

}  // namespace gp
