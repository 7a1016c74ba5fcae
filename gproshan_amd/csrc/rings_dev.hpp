// Ring builder device code (BFS with discovery-key ordering), shared by the
// toplesets kernel and the batched solver.  See rings.hip for the method.
#pragma once

#include "device.hpp"

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
__device__ inline uint32_t block_exclusive_scan(uint32_t x, uint32_t *s_w, uint32_t &total)
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

struct BfsJob {
	const uint32_t *src;   // sources (distinct)
	uint32_t ns;
	uint32_t *level;       // [nv]
	unsigned long long *key;
	uint32_t *sorted;      // [nv]
	uint32_t *limits;      // [nv + 2]
	uint32_t *n_limits;
};

// One BFS run by the whole workgroup (blockDim.x threads, a multiple of 64).
__device__ inline void bfs_block(const uint32_t *fan_off, const uint32_t *fan, const uint8_t *closed, const BfsJob &J)
{
	__shared__ uint32_t s_w[17];
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
	__syncthreads();
}



}  // namespace gp
