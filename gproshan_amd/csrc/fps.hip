// Farthest-point sampling over GPU PTP (farthest_point_sampling_ptp_gpu,
// src/cuda/geodesics_ptp.cu:87-172), device-resident.
//
// Each round runs the reference's multi-source GPU driver on the current
// samples -- rings from the samples (che::compute_toplesets), the window
// schedule through sorted[] (src/cuda/geodesics_ptp.cu:174-232, last-written
// buffer) -- takes the first vertex of largest distance (cuBLAS I?amax
// semantics: smallest index among the maxima) and appends it.  Here the
// samples, rings, schedule state and distances stay in HBM, and the host only
// polls done flags: no per-round upload, download or host BFS.
//
// Rings: the BFS levels from the sample set are kept from round to round; a
// new sample only lowers the levels of the vertices nearer to it, so each
// round runs a frontier BFS from the new sample alone (one launch per level,
// all workgroups, replayed from graphs) that stops where the old
// levels are already as low, then a counting sort by level rebuilds sorted[]
// and limits.  The order inside a ring differs from the reference's discovery
// order; a sweep relaxes a whole window of rings from the previous buffer, so
// the order inside a ring changes neither any value nor any ring test, and
// the samples and radius are the reference's.  Each round gathers the mesh's
// vertex-major slot tables into ring-ordered chunks (positions: ring 0 first,
// neighbour ids mapped to positions) so that the sweeps read coalesced slot
// rows as the single-field sweep kernel does; the maximum is taken over the
// vertices in their own numbering through the positions.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <type_traits>
#include <vector>

#include "device.hpp"
#include "levels.hpp"
#include "rings_dev.hpp"
#include "sweep_dev.hpp"

namespace gp {

struct FSlot {
	uint32_t i, j, d, iter, done, fresh, pad0, pad1;
};
struct FState {
	FSlot slot[4];
	uint32_t bad[4];
	uint32_t done, sweeps, final_d, nl;
	unsigned long long best;   // ordered bits of the round's largest distance
	uint32_t arg, stop, nr, pad3;         // nr: vertices reached this round
};

// Tables of a round: the rings of the current samples, gathered into
// ring-ordered chunks (positions), read by relax_lane.
struct FTables {
	const uint32_t *cid;
	const float2 *cgeo;
	GeoRemap64 cgeo64;
	const Edge64 *cx64;
	float h, inv_h;
};

struct FArgs {
	FTables t;
	const uint32_t *coff;                   // slot rows of chunk c: coff[c] .. coff[c + 1]
	const uint32_t *limits, *nl;            // rings of the current samples (device)
	void *dist0, *dist1;                    // by position
	FState *st;
	double tol;
};

// Sweep k of a round: the reference loop tail on the previous state, then the
// window relaxation as the single-field sweep kernel does it (one wave per
// chunk of 64 positions, XCD-aware split), with |limits| read from the device
// because the rings are rebuilt every round.
template <bool F64, uint32_t MS>
__global__ __launch_bounds__(256) void k_fps_sweep(FArgs a, uint32_t k4)
{
	using T = typename std::conditional<F64, double, float>::type;
	FState *st = a.st;
	const uint32_t nl = *a.nl, max_iter = nl << 1;
	const uint32_t kp = (k4 + 3u) & 3u;
	const FSlot prev = st->slot[kp];
	const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
	if (prev.done) {
		if (lead) st->slot[k4] = prev;
		return;
	}
	FSlot cur = prev;
	cur.fresh = 0;
	if (!prev.fresh) {
		const uint32_t nbad = st->bad[kp];
		const uint32_t pi = prev.i < (prev.j >> 1) ? (prev.j >> 1) : prev.i;
		cur.i = pi + (nbad == 0 ? 1u : 0u);
		cur.j = prev.j + (prev.j < nl - 1 ? 1u : 0u);
		cur.d = prev.d ^ 1u;
		cur.iter = prev.iter + 1;
		cur.done = !(cur.i < cur.j) || !(cur.iter < max_iter);
		if (lead) st->sweeps = cur.iter;
	}
	if (lead) {
		st->slot[k4] = cur;
		st->bad[(k4 + 1u) & 3u] = 0;
		if (cur.done) { st->final_d = cur.d; st->done = 1; }
	}
	if (cur.done) return;
	const uint32_t i = cur.i < (cur.j >> 1) ? (cur.j >> 1) : cur.i;   // K/2 limit band size
	const uint32_t start = a.limits[i], cend = a.limits[i + 1], end = a.limits[cur.j];
	const T *od = (const T *) (cur.d ? a.dist1 : a.dist0);
	T *nd = (T *) (cur.d ? a.dist0 : a.dist1);
	const uint32_t c0 = start >> 6, c1 = (end + 63u) >> 6;
	const uint32_t lane = threadIdx.x & 63u;
	// XCD-aware split: workgroups b % 8 == x take the x-th contiguous eighth
	// of the window's chunks.
	const uint32_t ngrp = gridDim.x < 8u ? gridDim.x : 8u;
	const uint32_t x = blockIdx.x & 7u;
	const uint32_t nbx = (gridDim.x - x + 7u) >> 3;
	const uint32_t wpb = blockDim.x >> 6;
	const uint32_t lw = (blockIdx.x >> 3) * wpb + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);   // wave-uniform
	const uint32_t part = (c1 - c0 + ngrp - 1) / ngrp;
	const uint32_t lo = c0 + x * part, hi = min(c1, lo + part);
	uint32_t bad = 0;
	for (uint32_t c = lo + lw; c < hi; c += nbx * wpb) {
		const uint32_t pos = (c << 6) + lane;
		bool lbad = false;
		if (pos >= start && pos < end) {
			const T old = od[pos];
			const uint32_t cb = a.coff[c];
			const T best = relax_lane<F64, 64, false, false, MS>(a.t, (size_t) cb * 64 + lane, a.coff[c + 1] - cb, od, old);
			nd[pos] = best;
			if (pos < cend) lbad = !(fabs((double) best - (double) old) / (double) old < a.tol);
		}
		bad += (uint32_t) __popcll(__ballot(lbad));
	}
	if (lane == 0 && bad) atomicAdd(&st->bad[k4], bad);
}

// Positions of a round: inv[v] = p for the reached vertices (sorted[p] = v),
// NIL for the others.
__global__ void k_fps_inv_clear(uint32_t *inv, uint32_t n)
{
	for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) inv[v] = GP_NIL;
}
__global__ void k_fps_inv(const uint32_t *sorted, const FState *st, uint32_t *inv)
{
	const uint32_t nr = st->nr;
	for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < nr; p += gridDim.x * blockDim.x) inv[sorted[p]] = p;
}

// Round set-up: both distance buffers at infinity.
template <typename T>
__global__ void k_fps_prep(T *d0, T *d1, uint32_t n)
{
	for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
		d0[v] = (T) INFINITY; d1[v] = (T) INFINITY;
	}
}

// The samples are ring 0: positions 0 .. ns - 1.
template <typename T>
__global__ void k_fps_sources(T *d0, T *d1, const BfsJob *job)
{
	for (uint32_t q = threadIdx.x; q < job->ns; q += blockDim.x) { d0[q] = (T) 0; d1[q] = (T) 0; }
}

__global__ void k_fps_state_init(FState *st, const uint32_t *nl, const uint32_t *limits)
{
	uint32_t *w = (uint32_t *) st;
	for (uint32_t q = 0; q < sizeof(FState) / 4; q++) w[q] = 0;
	const uint32_t n = *nl;
	st->nl = n;
	st->nr = limits[n - 1];
	st->slot[3].i = 1; st->slot[3].j = 2; st->slot[3].fresh = 1;
	st->slot[3].done = n < 3 ? 1u : 0u;     // only the sources: no sweep
	st->done = st->slot[3].done;
}

__device__ __forceinline__ unsigned long long ordered_bits(float x) { return __float_as_uint(x); }
__device__ __forceinline__ unsigned long long ordered_bits(double x) { return (unsigned long long) __double_as_longlong(x); }

// Distance of vertex v in the buffer d (by position); unreached vertices are
// +inf, as in the reference's field.
template <typename T>
__device__ __forceinline__ T vertex_dist(const T *d, const uint32_t *inv, uint32_t v)
{
	return inv[v] != GP_NIL ? d[inv[v]] : (T) INFINITY;
}

// Largest |distance| of the buffer the last sweep wrote, read over the
// positions (contiguous), +inf when fewer than n vertices were reached
// (distances are >= 0 or +inf, so their IEEE bit patterns order like the
// values).
template <typename T>
__global__ void k_fps_max(const void *d0, const void *d1, uint32_t n, FState *st)
{
	const T *d = (const T *) (st->final_d ? d1 : d0);
	const uint32_t nr = st->nr;
	unsigned long long m = 0;
	for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < nr; p += gridDim.x * blockDim.x) {
		const unsigned long long b = ordered_bits(fabs(d[p]));
		m = b > m ? b : m;
	}
	if (nr < n) m = ordered_bits((T) INFINITY);
	for (uint32_t o = 32; o > 0; o >>= 1) {
		const unsigned long long y = __shfl_xor(m, o);
		m = y > m ? y : m;
	}
	if ((threadIdx.x & 63u) == 0) atomicMax(&st->best, m);
}

// First vertex holding that maximum (the smallest index, as I?amax returns).
template <typename T>
__global__ void k_fps_arg(const void *d0, const void *d1, const uint32_t *inv, uint32_t n, FState *st)
{
	const T *d = (const T *) (st->final_d ? d1 : d0);
	const unsigned long long best = st->best;
	uint32_t a = GP_NIL;
	for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x)
		if (ordered_bits(fabs(vertex_dist(d, inv, v))) == best) { a = v; break; }
	for (uint32_t o = 32; o > 0; o >>= 1) a = min(a, (uint32_t) __shfl_xor(a, o));
	if ((threadIdx.x & 63u) == 0 && a != GP_NIL) atomicMin(&st->arg, a);
}

__global__ void k_fps_arg_init(FState *st) { st->arg = GP_NIL; st->best = 0; }

// Appends the new sample.  A maximum of 0 means every vertex is already a
// sample (the reference would append a duplicate and overrun its rings): stop.
template <typename T>
__global__ void k_fps_append(BfsJob *job, uint32_t *samples, FState *st)
{
	const unsigned long long zero = ordered_bits((T) 0);
	if (st->best == zero || st->arg == GP_NIL) { st->stop = 1; return; }
	samples[job->ns] = st->arg;
	job->ns = job->ns + 1;
}

// ---------------------------------------------------------------------------

struct FpsRun {
	uint32_t *cid = nullptr, *midx = nullptr, *inv = nullptr, *samples = nullptr;
	float2 *cgeo = nullptr;
	Edge64 *cx64 = nullptr;
	void *d0 = nullptr, *d1 = nullptr;
	FState *st = nullptr, *h_st = nullptr;   // h_st: the round's snapshot, then two sweep snapshots
	BfsJob *job = nullptr;
	hipStream_t stream = nullptr;
	hipEvent_t e0 = nullptr, e1 = nullptr, ev[2] = { nullptr, nullptr };
	hipGraphExec_t g_big = nullptr, g_small = nullptr;
	~FpsRun()
	{
		for (void *p : { (void *) cid, (void *) midx, (void *) inv, (void *) samples, (void *) cgeo, (void *) cx64,
		                 d0, d1, (void *) st, (void *) job })
			(void) hipFree(p);
		if (h_st) (void) hipHostFree(h_st);
		if (g_big) (void) hipGraphExecDestroy(g_big);
		if (g_small) (void) hipGraphExecDestroy(g_small);
		if (e0) (void) hipEventDestroy(e0);
		if (e1) (void) hipEventDestroy(e1);
		for (hipEvent_t e : ev)
			if (e) (void) hipEventDestroy(e);
		if (stream) (void) hipStreamDestroy(stream);
	}
};

template <typename V>
static int fup(V *&dst, const V *src, size_t n)
{
	if (hipMalloc(&dst, sizeof(V) * std::max<size_t>(n, 1)) != hipSuccess) return fail(GP_ERR_NOMEM, "fps: allocation failed");
	if (n) GP_HIP(hipMemcpy(dst, src, sizeof(V) * n, hipMemcpyHostToDevice));
	return GP_OK;
}

template <bool F64>
static int fps_impl(Che &m, uint32_t *samples, uint32_t n_init, uint32_t cap, uint64_t n, double radio,
                    uint32_t *n_out, double *radio_out, double *seconds)
{
	using T = typename std::conditional<F64, double, float>::type;
	const uint32_t nv = m.nv;
	for (uint32_t v = 0; v < nv; v++)
		if (m.fan_off[v + 1] - m.fan_off[v] > MAX_LINK) return fail(GP_ERR_CAPACITY, "fps: vertex %u: link over %u", v, MAX_LINK);
	// The reference's sample budget (src/cuda/geodesics_ptp.cu:123-126).
	if (n >= nv) n = nv >> 1;
	const uint64_t rounds = n > n_init ? n - n_init : 0;
	if (n_init + rounds > cap) return fail(GP_ERR_CAPACITY, "fps: samples needs room for %llu entries",
	                                       (unsigned long long) (n_init + rounds));
	VertexTables mt{};
	uint32_t S = 0;
	int rc = master_tables(m, F64, mt, S);
	if (rc) return rc;
	FpsRun r;
	// Ring-ordered chunked tables of a round: each chunk as many slot rows as
	// its longest fan (each ring sorted longest fans first), room for the
	// longest fan in every chunk; fans past the register-held slots take the
	// streaming loop.
	const uint32_t nchunks = (nv + 63) / 64, npos = nchunks * 64;
	const size_t lanes = (size_t) nchunks * S * 64;
	if (hipMalloc(&r.cid, 4 * lanes) != hipSuccess || hipMalloc(&r.inv, 4ull * nv) != hipSuccess ||
	    (F64 ? hipMalloc(&r.cx64, sizeof(Edge64) * lanes)
	         : (hipMalloc(&r.cgeo, sizeof(float2) * lanes) == hipSuccess ? hipMalloc(&r.midx, 4 * lanes)
	                                                                      : hipErrorOutOfMemory)) != hipSuccess)
		return fail(GP_ERR_NOMEM, "fps: table allocation failed");
	if ((rc = fup(r.samples, samples, cap))) return rc;
	if (hipMalloc(&r.d0, sizeof(T) * nv) != hipSuccess || hipMalloc(&r.d1, sizeof(T) * nv) != hipSuccess ||
	    hipMalloc(&r.st, sizeof(FState)) != hipSuccess || hipMalloc(&r.job, sizeof(BfsJob)) != hipSuccess ||
	    hipHostMalloc(&r.h_st, 3 * sizeof(FState)) != hipSuccess)
		return fail(GP_ERR_NOMEM, "fps: device allocation failed");
	GP_HIP(hipStreamCreateWithFlags(&r.stream, hipStreamNonBlocking));
	Levels lv;
	if ((rc = lv.create(m.dev, r.stream))) return rc;
	const BfsJob job{ r.samples, n_init, lv.level, nullptr, lv.sorted, lv.limits, lv.nl };
	GP_HIP(hipMemcpy(r.job, &job, sizeof(job), hipMemcpyHostToDevice));
	GP_HIP(hipEventCreate(&r.e0));
	GP_HIP(hipEventCreate(&r.e1));
	GP_HIP(hipEventCreateWithFlags(&r.ev[0], hipEventDisableTiming));
	GP_HIP(hipEventCreateWithFlags(&r.ev[1], hipEventDisableTiming));
	FArgs a{};
	a.t.cid = r.cid; a.t.cgeo = r.cgeo; a.t.cgeo64 = GeoRemap64{ mt.cgeo64, r.midx }; a.t.cx64 = r.cx64;
	a.t.h = m.dev->h; a.t.inv_h = m.dev->inv_h;
	a.coff = lv.coff; a.limits = lv.limits; a.nl = lv.nl; a.dist0 = r.d0; a.dist1 = r.d1; a.st = r.st;
	a.tol = GP_PTP_TOL;
	int dev = 0, ncu = 0, occ = 0;
	GP_HIP(hipGetDevice(&dev));
	GP_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
	const bool ms8 = S <= 8;
	// One wave per workgroup, as the single-field sweeps.
	constexpr uint32_t FB = 64;
	GP_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, ms8 ? k_fps_sweep<F64, 8> : k_fps_sweep<F64, MAXS>, FB, 0));
	const uint32_t grid = std::max<uint32_t>(8, std::min<uint32_t>((uint32_t) ncu * (uint32_t) std::max(1, occ),
	                                                               nchunks));
	constexpr uint32_t BIG = 64, SMALL = 8;
	auto capture = [&](uint32_t count, hipGraphExec_t *out) -> int {
		FArgs fa = a;
		const uint32_t k4s[4] = { 0, 1, 2, 3 };
		const void *fn = ms8 ? (const void *) k_fps_sweep<F64, 8> : (const void *) k_fps_sweep<F64, MAXS>;
		GraphChain gc;
		for (uint32_t q = 0; q < count; q++) {
			void *args[] = { &fa, (void *) &k4s[q & 3u] };
			gc.add(fn, dim3(grid), dim3(FB), args);
		}
		const hipError_t e = gc.instantiate(out);
		if (e != hipSuccess) return fail(GP_ERR_HIP, "fps graph: %s", hipGetErrorString(e));
		return GP_OK;
	};
	const bool graphs = graphs_enabled();
	if (graphs && ((rc = capture(BIG, &r.g_big)) || (rc = capture(SMALL, &r.g_small)))) return rc;
	const uint32_t ig = std::max<uint32_t>(1, std::min<uint32_t>(2048, (nv + 255) / 256));
	const uint32_t gg = std::max<uint32_t>(1, std::min<uint32_t>(4096, (npos + 255) / 256));
	const auto t0 = std::chrono::steady_clock::now();
	GP_HIP(hipEventRecord(r.e0, r.stream));
	double max_dist = INFINITY;
	uint32_t count = n_init;
	for (uint64_t left = rounds; left-- > 0 && max_dist > radio;) {
		k_fps_prep<T><<<ig, 256, 0, r.stream>>>((T *) r.d0, (T *) r.d1, nv);
		k_fps_sources<T><<<1, 256, 0, r.stream>>>((T *) r.d0, (T *) r.d1, r.job);
		// Levels: from all initial samples in the first round, then lowered
		// from the newest sample.
		if (count == n_init) lv.clear();
		lv.seed(r.samples, count == n_init ? 0u : count - 1, count);
		if ((rc = lv.expand())) return rc;
		lv.sort();
		k_fps_state_init<<<1, 1, 0, r.stream>>>(r.st, lv.nl, lv.limits);
		k_fps_inv_clear<<<ig, 256, 0, r.stream>>>(r.inv, nv);
		k_fps_inv<<<ig, 256, 0, r.stream>>>(lv.sorted, r.st, r.inv);
		if (F64)
			k_gather_tables<true, false><<<gg, 256, 0, r.stream>>>(lv.sorted, r.inv, &r.st->nr, npos, S, lv.coff, mt, r.cid,
			                                                        nullptr, nullptr, nullptr, r.cx64);
		else
			k_gather_tables<false, true><<<gg, 256, 0, r.stream>>>(lv.sorted, r.inv, &r.st->nr, npos, S, lv.coff, mt, r.cid,
			                                                        r.cgeo, nullptr, r.midx, nullptr);
		snapshot(r.st, r.h_st, r.stream);
		GP_HIP(hipGetLastError());
		GP_HIP(hipStreamSynchronize(r.stream));
		const uint32_t nl = r.h_st->nl;
		if (nl < 2 || nl > nv + 2) return fail(GP_ERR_KERNEL, "fps: bad ring count %u", nl);
		// Sweeps: graph replays until the device reports done (at most
		// 2*|limits| sweeps plus the replay that records the last outcome),
		// two in flight: each replay is followed by a kernel writing the state
		// into its pinned snapshot and an event, and the host waits for the older
		// replay while the newer one runs (a replay queued behind the one that
		// finishes meets done slots and returns at once).  j grows by one ring
		// per sweep, so big graphs run while the last seen j, one replay
		// behind, leaves room for them.
		FState *snap = r.h_st + 1;
		uint32_t launched = 0, j_seen = 2;
		auto replay = [&](uint32_t s) -> int {
			const bool big = nl - 1 > j_seen + BIG + (launched ? BIG : 0);
			if (graphs) {
				GP_HIP(hipGraphLaunch(big ? r.g_big : r.g_small, r.stream));
			} else {
				for (uint32_t q = 0; q < (big ? BIG : SMALL); q++) {
					if (ms8) k_fps_sweep<F64, 8><<<grid, FB, 0, r.stream>>>(a, q & 3u);
					else k_fps_sweep<F64, MAXS><<<grid, FB, 0, r.stream>>>(a, q & 3u);
				}
			}
			launched += big ? BIG : SMALL;
			snapshot(r.st, snap + s, r.stream);
			GP_HIP(hipGetLastError());
			GP_HIP(hipEventRecord(r.ev[s], r.stream));
			return GP_OK;
		};
		if (!r.h_st->done) {
			if ((rc = replay(0))) return rc;
			for (uint32_t cur = 0;; cur ^= 1u) {
				if (launched > 2 * nl + 2 * BIG) return fail(GP_ERR_KERNEL, "fps: schedule did not finish");
				if ((rc = replay(cur ^ 1u))) return rc;
				GP_HIP(hipEventSynchronize(r.ev[cur]));
				if (snap[cur].done) break;
				j_seen = snap[cur].slot[3].j;
			}
		}
		k_fps_arg_init<<<1, 1, 0, r.stream>>>(r.st);
		k_fps_max<T><<<ig, 256, 0, r.stream>>>(r.d0, r.d1, nv, r.st);
		k_fps_arg<T><<<ig, 256, 0, r.stream>>>(r.d0, r.d1, r.inv, nv, r.st);
		k_fps_append<T><<<1, 1, 0, r.stream>>>(r.job, r.samples, r.st);
		snapshot(r.st, r.h_st, r.stream);
		GP_HIP(hipGetLastError());
		GP_HIP(hipStreamSynchronize(r.stream));
		if (r.h_st->stop) break;
		count++;
		// max_dist is read when a radius is given or in the last round
		// (src/cuda/geodesics_ptp.cu:145-146).
		if (radio > 0 || left == 0) {
			const unsigned long long b = r.h_st->best;
			if (F64) { double x; std::memcpy(&x, &b, 8); max_dist = x; }
			else { const uint32_t u = (uint32_t) b; float x; std::memcpy(&x, &u, 4); max_dist = x; }
		}
	}
	GP_HIP(hipEventRecord(r.e1, r.stream));
	GP_HIP(hipStreamSynchronize(r.stream));
	float ms = 0;
	GP_HIP(hipEventElapsedTime(&ms, r.e0, r.e1));
	GP_HIP(hipMemcpy(samples, r.samples, 4ull * count, hipMemcpyDeviceToHost));
	if (n_out) *n_out = count;
	if (radio_out) *radio_out = max_dist;
	if (seconds) *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
	(void) ms;
	return GP_OK;
}

int fps_gpu(Che &m, uint32_t *samples, uint32_t n_init, uint32_t cap, uint64_t n, double radio, bool f64,
            uint32_t *n_out, double *radio_out, double *seconds)
{
	if (!samples || !n_init) return fail(GP_ERR_INVALID, "fps: at least one initial sample is needed");
	std::vector<uint8_t> seen(m.nv, 0);
	for (uint32_t q = 0; q < n_init; q++) {
		if (samples[q] >= m.nv) return fail(GP_ERR_INVALID, "fps: sample %u out of range", samples[q]);
		if (seen[samples[q]]) return fail(GP_ERR_DUPLICATE_SOURCE, "fps: sample %u listed twice", samples[q]);
		seen[samples[q]] = 1;
	}
	return f64 ? fps_impl<true>(m, samples, n_init, cap, n, radio, n_out, radio_out, seconds)
	           : fps_impl<false>(m, samples, n_init, cap, n, radio, n_out, radio_out, seconds);
}

}  // namespace gp
