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

#include <algorithm>
#include <vector>

namespace gp {

constexpr uint32_t BFS_THREADS = 1024;
constexpr uint32_t MAX_LINK = 1024;

__device__ __forceinline__ uint32_t ld_relaxed(const uint32_t *p)
{
	return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned long long ld_relaxed(const unsigned long long *p)
{
	return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_relaxed(uint32_t *p, uint32_t v)
{
	__hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Exclusive prefix sum over the workgroup (wave64 shuffles + LDS for the
// per-wave totals).  All threads must call it.
__device__ uint32_t block_exclusive_scan(uint32_t x, uint32_t *s_w, uint32_t &total)
{
	const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
	uint32_t inc = x;
	for (uint32_t o = 1; o < 64; o <<= 1) {
		const uint32_t y = __shfl_up(inc, o);
		if (lane >= o) inc += y;
	}
	if (lane == 63) s_w[w] = inc;
	__syncthreads();
	if (threadIdx.x == 0) {
		uint32_t acc = 0;
		for (uint32_t q = 0; q < nw; q++) { const uint32_t t = s_w[q]; s_w[q] = acc; acc += t; }
		s_w[16] = acc;
	}
	__syncthreads();
	const uint32_t ex = inc - x + s_w[w];
	total = s_w[16];
	__syncthreads();
	return ex;
}

__global__ void k_rings_init(uint32_t *level, unsigned long long *key, uint32_t *sorted, uint32_t nv)
{
	for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < nv; v += gridDim.x * blockDim.x) {
		level[v] = GP_NIL;
		key[v] = ~0ull;
		sorted[v] = GP_NIL;
	}
}

struct BfsJob {
	const uint32_t *src;   // sources (distinct)
	uint32_t ns;
	uint32_t *level;       // [nv]
	unsigned long long *key;
	uint32_t *sorted;      // [nv]
	uint32_t *limits;      // [nv + 2]
	uint32_t *n_limits;
};

__global__ __launch_bounds__(BFS_THREADS) void k_rings_bfs(const uint32_t *fan_off, const uint32_t *fan,
                                                           const uint8_t *closed, const BfsJob *jobs)
{
	__shared__ uint32_t s_w[17];
	const BfsJob J = jobs[blockIdx.x];
	const uint32_t tid = threadIdx.x;
	for (uint32_t q = tid; q < J.ns; q += blockDim.x) {
		st_relaxed(&J.sorted[q], J.src[q]);
		st_relaxed(&J.level[J.src[q]], 0u);
	}
	if (tid == 0) J.limits[0] = 0;
	__syncthreads();

	uint32_t fb = 0, fe = J.ns, p = J.ns, nl = 1, lev = 0;
	while (fb < fe) {
		// Phase 1: smallest (pos, slot) key for every unreached link vertex.
		for (uint32_t pos = fb + tid; pos < fe; pos += blockDim.x) {
			const uint32_t v = ld_relaxed(&J.sorted[pos]);
			const uint32_t s = fan_off[v], c = fan_off[v + 1] - s;
			const uint32_t n = c ? (closed[v] ? c - 1 : c) : 0;
			for (uint32_t k = 0; k < n; k++) {
				const uint32_t u = fan[s + k] & ID_MASK;
				if (ld_relaxed(&J.level[u]) == GP_NIL)
					atomicMin(&J.key[u], ((unsigned long long) pos << 10) | k);
			}
		}
		__syncthreads();
		// Phase 2: owners emit the new ring in key order.
		uint32_t base = p;
		for (uint32_t c0 = fb; c0 < fe; c0 += blockDim.x) {
			const uint32_t pos = c0 + tid;
			uint32_t v = 0, s = 0, n = 0, cnt = 0;
			if (pos < fe) {
				v = ld_relaxed(&J.sorted[pos]);
				s = fan_off[v];
				const uint32_t c = fan_off[v + 1] - s;
				n = c ? (closed[v] ? c - 1 : c) : 0;
				for (uint32_t k = 0; k < n; k++) {
					const uint32_t u = fan[s + k] & ID_MASK;
					if (ld_relaxed(&J.level[u]) == GP_NIL &&
					    ld_relaxed(&J.key[u]) == (((unsigned long long) pos << 10) | k))
						cnt++;
				}
			}
			uint32_t total;
			uint32_t off = block_exclusive_scan(cnt, s_w, total);
			if (cnt) {
				uint32_t idx = base + off;
				for (uint32_t k = 0; k < n; k++) {
					const uint32_t u = fan[s + k] & ID_MASK;
					if (ld_relaxed(&J.level[u]) == GP_NIL &&
					    ld_relaxed(&J.key[u]) == (((unsigned long long) pos << 10) | k))
						st_relaxed(&J.sorted[idx++], u);
				}
			}
			base += total;
		}
		__syncthreads();
		// Phase 3: the new ring gets its level.
		for (uint32_t q = p + tid; q < base; q += blockDim.x)
			st_relaxed(&J.level[ld_relaxed(&J.sorted[q])], lev + 1);
		__syncthreads();
		if (base > p) {
			if (tid == 0) J.limits[nl] = p;
			nl++;
		}
		fb = p; fe = base; p = base; lev++;
	}
	if (tid == 0) {
		J.limits[nl] = p;
		*J.n_limits = nl + 1;
	}
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
