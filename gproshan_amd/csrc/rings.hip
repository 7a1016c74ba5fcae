// Toplesets (BFS rings) on the GPU, bit-identical to che::compute_toplesets
// (src/che.cpp:546-593).
//
// Ring membership is the hop distance, which any level-synchronous BFS
// reproduces.  The order inside a ring is the reference's discovery order:
// the serial BFS appends u when it first meets it while scanning the ring
// before in sorted order and, within a vertex, in link order.  So each newly
// reached u takes the smallest key (position of the discovering vertex in
// sorted) * 1024 + (slot in its link) over all its discoverers
// (64-bit atomicMin), and the new ring is written in key order by letting the
// owner (pos, slot) of each u emit it, with a workgroup prefix sum over the
// frontier in (pos, slot) order.  One 1024-thread workgroup runs one BFS, so
// many BFS jobs can share a launch.
#include "device.hpp"
#include "rings_dev.hpp"

#include <algorithm>
#include <vector>

namespace gp {

__global__ void k_rings_init(uint32_t *level, unsigned long long *key, uint32_t *sorted, uint32_t nv)
{
	for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < nv; v += gridDim.x * blockDim.x) {
		level[v] = GP_NIL;
		key[v] = ~0ull;
		sorted[v] = GP_NIL;
	}
}

__global__ __launch_bounds__(BFS_THREADS) void k_rings_bfs(const uint32_t *fan_off, const uint32_t *fan,
                                                           const uint8_t *closed, const BfsJob *jobs)
{
	bfs_block(fan_off, fan, closed, jobs[blockIdx.x]);
}

int toplesets_gpu(Che &m, const uint32_t *src, uint32_t ns, uint32_t *levels, uint32_t *sorted,
                  std::vector<uint32_t> &limits)
{
	limits.clear();
	if (ns == 0) return GP_OK;
	std::vector<uint8_t> seen(m.nv, 0);
	for (uint32_t q = 0; q < ns; q++) {
		if (src[q] >= m.nv) return fail(GP_ERR_INVALID, "toplesets: source %u out of range", src[q]);
		if (seen[src[q]]) return fail(GP_ERR_DUPLICATE_SOURCE, "toplesets: source %u listed twice", src[q]);
		seen[src[q]] = 1;
	}
	for (uint32_t v = 0; v < m.nv; v++)
		if (m.fan_off[v + 1] - m.fan_off[v] > MAX_LINK) return fail(GP_ERR_CAPACITY, "vertex %u: link over %u", v, MAX_LINK);
	int rc = upload_mesh(m);
	if (rc) return rc;
	const DeviceMesh *dm = m.dev;
	const uint32_t nv = m.nv;
	uint32_t *d_level = nullptr, *d_sorted = nullptr, *d_limits = nullptr, *d_nl = nullptr, *d_src = nullptr;
	unsigned long long *d_key = nullptr;
	BfsJob *d_job = nullptr;
	struct Free {
		std::vector<void *> p;
		~Free() { for (void *q : p) (void) hipFree(q); }
	} fr;
	if (hipMalloc(&d_level, 4ull * nv) != hipSuccess || hipMalloc(&d_sorted, 4ull * nv) != hipSuccess ||
	    hipMalloc(&d_limits, 4ull * (nv + 2)) != hipSuccess || hipMalloc(&d_nl, 4) != hipSuccess ||
	    hipMalloc(&d_src, 4ull * ns) != hipSuccess || hipMalloc(&d_key, 8ull * nv) != hipSuccess ||
	    hipMalloc(&d_job, sizeof(BfsJob)) != hipSuccess) {
		fr.p = { d_level, d_sorted, d_limits, d_nl, d_src, d_key, d_job };
		return fail(GP_ERR_NOMEM, "toplesets_gpu: allocation failed");
	}
	fr.p = { d_level, d_sorted, d_limits, d_nl, d_src, d_key, d_job };
	BfsJob job{ d_src, ns, d_level, d_key, d_sorted, d_limits, d_nl };
	GP_HIP(hipMemcpy(d_src, src, 4ull * ns, hipMemcpyHostToDevice));
	GP_HIP(hipMemcpy(d_job, &job, sizeof(job), hipMemcpyHostToDevice));
	const uint32_t g = std::max<uint32_t>(1, std::min<uint32_t>(2048, (nv + 255) / 256));
	k_rings_init<<<g, 256>>>(d_level, d_key, d_sorted, nv);
	GP_HIP(hipGetLastError());
	k_rings_bfs<<<1, BFS_THREADS>>>(dm->fan_off, dm->fan, dm->closed, d_job);
	GP_HIP(hipGetLastError());
	GP_HIP(hipDeviceSynchronize());
	uint32_t nl = 0;
	GP_HIP(hipMemcpy(&nl, d_nl, 4, hipMemcpyDeviceToHost));
	if (nl < 2 || nl > nv + 2) return fail(GP_ERR_KERNEL, "toplesets_gpu: bad ring count %u", nl);
	limits.resize(nl);
	GP_HIP(hipMemcpy(limits.data(), d_limits, 4ull * nl, hipMemcpyDeviceToHost));
	GP_HIP(hipMemcpy(levels, d_level, 4ull * nv, hipMemcpyDeviceToHost));
	GP_HIP(hipMemcpy(sorted, d_sorted, 4ull * nv, hipMemcpyDeviceToHost));
	return GP_OK;
}

}  // namespace gp
