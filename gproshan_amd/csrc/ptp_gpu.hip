// gfx950 parallel toplesets propagation: device mesh tables, the window sweep
// kernels and the device-resident solver plan.
//
// Schedule (src/geodesics_ptp.cpp:61-109; src/cuda/geodesics_ptp.cu:198-231):
// rings [i, j) form the window, i is raised to j/2 when it lags (K/2 band
// limit), ring i retires when every vertex in it changed by less than tol
// (relative), j grows by one ring per sweep, the double buffer flips each
// sweep, and at most 2*|limits| sweeps run.
//
// The reference launches a relax kernel, a relative-error kernel and a Thrust
// count per sweep with three host synchronisations.  Here the schedule lives
// in device memory: a sweep kernel relaxes the window it reads from there and
// counts non-converged ring-i vertices (wave reductions, one atomic per wave
// with a non-zero count), and a one-thread kernel advances the schedule.  The
// pairs are captured once into a hipGraph and replayed until the device sets
// its done flag; the host polls that flag once per replay.  (A persistent
// kernel with a grid barrier per sweep measured slower on MI355X: the
// agent-scope fence each workgroup needs to publish its writes across the
// eight XCD L2s costs ~36 ns per workgroup, ~19 us per barrier at 512
// workgroups, against ~1.6 us per graph-launched kernel.)
//
// Work layout (see Layout in gp_internal.hpp): ELL, one lane per vertex in
// ring order, 64 vertices per wave.  A wave issues all slot loads of its
// chunk (fan ids and triangle geometry, one contiguous access per slot), then
// all neighbour-distance gathers, then runs the triangle updates from
// registers: two dependent memory round trips per 64 vertices.  Each XCD
// takes one contiguous eighth of the window's chunks so neighbouring gathers
// share its L2.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <type_traits>

#include "device.hpp"
#include "sweep_dev.hpp"
#include "update_step.hpp"

namespace gp {

void device_mesh_free(DeviceMesh *d)
{
	if (!d) return;
	(void) hipFree(d->fan_off);
	(void) hipFree(d->fan);
	(void) hipFree(d->pos);
	(void) hipFree(d->closed);
	delete d;
}

int device_check(char *name, size_t cap)
{
	int n = 0;
	if (hipGetDeviceCount(&n) != hipSuccess || n == 0)
		return fail(GP_ERR_NO_DEVICE, "no HIP device");
	hipDeviceProp_t prop;
	GP_HIP(hipGetDeviceProperties(&prop, 0));
	if (name && cap) {
		std::strncpy(name, prop.gcnArchName, cap - 1);
		name[cap - 1] = 0;
	}
	if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
		return fail(GP_ERR_NO_DEVICE, "device is %s, this library is built for gfx950", prop.gcnArchName);
	return GP_OK;
}

// Uploads the fan tables (ring builder) and FP64 positions once per mesh.
int upload_mesh(Che &m)
{
	if (m.dev) return GP_OK;
	int rc = device_check(nullptr, 0);
	if (rc) return rc;
	const uint32_t nv = m.nv, nfan = m.fan_off[nv];
	auto dm = std::make_unique<DeviceMesh>();
	dm->nv = nv; dm->nfan = nfan;
	dm->h = (float) m.mean_edge; dm->inv_h = (float) (1.0 / m.mean_edge);
	auto cleanup = [&](int r) { device_mesh_free(dm.release()); return r; };
	if (hipMalloc(&dm->fan_off, sizeof(uint32_t) * (nv + 1)) != hipSuccess ||
	    hipMalloc(&dm->fan, sizeof(uint32_t) * std::max<uint32_t>(nfan, 1)) != hipSuccess ||
	    hipMalloc(&dm->pos, sizeof(double) * 3 * (size_t) nv) != hipSuccess ||
	    hipMalloc(&dm->closed, nv) != hipSuccess)
		return cleanup(fail(GP_ERR_NOMEM, "device mesh allocation failed"));
	if (hipMemcpy(dm->fan_off, m.fan_off.data(), sizeof(uint32_t) * (nv + 1), hipMemcpyHostToDevice) != hipSuccess ||
	    (nfan && hipMemcpy(dm->fan, m.fan.data(), sizeof(uint32_t) * nfan, hipMemcpyHostToDevice) != hipSuccess) ||
	    hipMemcpy(dm->pos, m.gt.data(), sizeof(double) * 3 * (size_t) nv, hipMemcpyHostToDevice) != hipSuccess ||
	    hipMemcpy(dm->closed, m.fan_closed.data(), nv, hipMemcpyHostToDevice) != hipSuccess)
		return cleanup(fail(GP_ERR_HIP, "device mesh upload failed"));
	m.dev = dm.release();
	return GP_OK;
}

// ---------------------------------------------------------------------------
// Kernels

// Schedule state in device memory.  Sweep kernel k of a run (k counted from
// 0) reads the state of sweep k-1 from slot[(k-1) & 3] together with its
// non-converged count bad[(k-1) & 3], advances it (i, j, d as the reference
// loop does), publishes it in slot[k & 3] (workgroup 0), zeroes bad[(k+1) & 3]
// for the next sweep, and relaxes its own window counting into bad[k & 3].
// No workgroup reads a slot or counter while another workgroup of the same
// launch writes it, so kernel boundaries are the only synchronisation.  The
// captured graphs hold a multiple of 4 sweeps so k & 3 stays continuous.
struct Slot {
	uint32_t i, j, d, iter, done, fresh, phase, judge;
};
struct PState {
	Slot slot[4];
	uint32_t bad[4];
	uint32_t done, sweeps, final_d, pad;
	unsigned long long relax, cond, f64;
};

struct KArgs {
	const uint32_t *limits;
	const uint32_t *coff;      // per chunk: first slot row
	const uint32_t *cid;
	const float4 *cgeo;
	const double4 *cgeo64;
	const double4 *cx64;
	const unsigned long long *fcum;  // FP64-update triangles of positions [0, pos)
	const uint32_t *order;     // position -> vertex; null when the mesh is numbered by rings
	void *dist0, *dist1;
	PState *st;
	uint32_t *trace;           // 8 words per sweep, or null
	const uint16_t *lidx;      // band solves: slot neighbour's index in its tile, or LOCAL_OUT
	double tol;
	float h, inv_h;
	uint32_t nl, max_iter;
	uint32_t nr, local_iters, judge_cap;
};

// Relaxes the window [start, end) (positions); returns this thread's count of
// non-converged vertices among positions [start, cend).
template <bool F64>
__device__ __forceinline__ uint32_t sweep_body(const KArgs &a, uint32_t start, uint32_t end, uint32_t cend,
                                               const typename std::conditional<F64, double, float>::type *od,
                                               typename std::conditional<F64, double, float>::type *nd)
{
	using T = typename std::conditional<F64, double, float>::type;
	if (start >= end) return 0;
	const uint32_t c0 = start >> 6, c1 = (end + 63u) >> 6;
	const uint32_t lane = threadIdx.x & 63u;
	// XCD-aware split: workgroups b with b % 8 == x run on XCD x and take the
	// x-th contiguous eighth of the window's chunks.
	const uint32_t ngrp = gridDim.x < 8u ? gridDim.x : 8u;
	const uint32_t x = blockIdx.x & 7u;
	const uint32_t nbx = (gridDim.x - x + 7u) >> 3;
	const uint32_t wpb = blockDim.x >> 6;
	const uint32_t lw = (blockIdx.x >> 3) * wpb + (threadIdx.x >> 6);
	const uint32_t nw_x = nbx * wpb;
	const uint32_t len = c1 - c0;
	const uint32_t part = (len + ngrp - 1) / ngrp;
	const uint32_t lo = c0 + x * part;
	const uint32_t hi = min(c1, lo + part);
	uint32_t bad = 0;
	for (uint32_t c = lo + lw; c < hi; c += nw_x) {
		const uint32_t pos = (c << 6) + lane;
		const uint32_t cb = a.coff[c], S = a.coff[c + 1] - cb;
		if (pos >= start && pos < end) {
			const uint32_t v = a.order ? a.order[pos] : pos;
			const T old = od[v];
			const T best = relax_lane<F64>(a, (size_t) cb * 64 + lane, S, od, old);
			nd[v] = best;
			if (pos < cend) bad += !(fabs((double) best - (double) old) / (double) old < a.tol);
		}
	}
	return bad;
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t x)
{
	for (uint32_t o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
	return x;
}

// Sweep k (k4 = k & 3) of a run: advance the schedule, then relax its window.
template <bool F64>
__global__ __launch_bounds__(256) void k_sweep(KArgs a, uint32_t k4)
{
	using T = typename std::conditional<F64, double, float>::type;
	PState *st = a.st;
	const uint32_t kp = (k4 + 3u) & 3u;
	const Slot prev = st->slot[kp];
	const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
	if (prev.done) {
		// Later launches of the replay read slot k & 3 in turn: keep it done.
		if (lead) st->slot[k4] = prev;
		return;
	}
	Slot cur = prev;
	cur.fresh = 0;
	if (!prev.fresh) {
		// The previous sweep's outcome: the reference loop tail.
		const uint32_t nbad = st->bad[kp];
		const uint32_t pi = prev.i < (prev.j >> 1) ? (prev.j >> 1) : prev.i;
		cur.i = pi + (nbad == 0 ? 1u : 0u);
		cur.j = prev.j + (prev.j < a.nl - 1 ? 1u : 0u);
		cur.d = prev.d ^ 1u;
		cur.iter = prev.iter + 1;
		cur.done = !(cur.i < cur.j) || !(cur.iter < a.max_iter);
		if (lead) {
			const uint32_t l0 = a.limits[pi], l1 = a.limits[pi + 1], lj = a.limits[prev.j];
			const uint32_t ncond = l1 - l0;
			if (a.trace) {
				uint32_t *r = a.trace + 8 * (size_t) prev.iter;
				r[0] = prev.iter + 1; r[1] = pi; r[2] = prev.j; r[3] = l0; r[4] = lj; r[5] = ncond;
				r[6] = ncond - nbad; r[7] = 0;
			}
			st->relax += lj - l0;
			st->cond += ncond;
			if (a.fcum) st->f64 += a.fcum[lj] - a.fcum[l0];
			st->sweeps = cur.iter;
		}
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
	const uint32_t bad = wave_sum(sweep_body<F64>(a, start, end, cend, od, nd));
	if ((threadIdx.x & 63u) == 0 && bad) atomicAdd(&st->bad[k4], bad);
}

// ---------------------------------------------------------------------------
// Band schedule (learned / Euclidean / ring-group bands; no reference
// counterpart).  Vertices are numbered by band, then by Morton code inside a
// band, so a tile of TILE consecutive positions is a compact patch of one or
// two bands.  A workgroup loads its tile's distances into LDS and runs up to
// local_iters Jacobi iterations on it (out-of-tile neighbours held at their
// values from the previous sweep), so the front crosses up to local_iters
// hops per sweep instead of one.  The window [i, j) of bands grows by one band
// per sweep; band i retires when it changed by less than tol and band i+1 was
// relaxed with it.  When the window closes, judge sweeps relax the whole mesh
// until no vertex changes by tol: a bad prediction costs sweeps, never
// accuracy.

constexpr uint32_t TILE = 256;

__device__ __forceinline__ float lb_gather(uint32_t li, const float *sd, float tout)
{
	return li != LOCAL_OUT ? sd[li] : tout;
}

// Relaxes positions [ws, we) tile by tile; returns this thread's count of
// non-converged vertices among [cs, ce).
__device__ uint32_t lb_tiles(const KArgs &a, uint32_t ws, uint32_t we, uint32_t cs, uint32_t ce, const float *od,
                             float *nd)
{
	__shared__ float sd[2][TILE];
	if (ws >= we) return 0;
	const uint32_t t0 = ws / TILE, t1 = (we + TILE - 1) / TILE;
	const uint32_t ngrp = gridDim.x < 8u ? gridDim.x : 8u;
	const uint32_t x = blockIdx.x & 7u;
	const uint32_t nbx = (gridDim.x - x + 7u) >> 3;
	const uint32_t len = t1 - t0;
	const uint32_t part = (len + ngrp - 1) / ngrp;
	const uint32_t lo = t0 + x * part;
	const uint32_t hi = min(t1, lo + part);
	const uint32_t tid = threadIdx.x, lane = tid & 63u;
	uint32_t bad = 0;
	for (uint32_t tile = lo + (blockIdx.x >> 3); tile < hi; tile += nbx) {
		const uint32_t pos = tile * TILE + tid;
		const bool live = pos < a.nr;
		const bool act = live && pos >= ws && pos < we;
		uint32_t S = 0;
		size_t base = 0;
		if (live) {
			const uint32_t c = pos >> 6, cb = a.coff[c];
			S = a.coff[c + 1] - cb;
			base = (size_t) cb * 64 + lane;
		}
		const float old = live ? od[pos] : (float) INFINITY;
		const bool fast = S <= MAXS;
		uint32_t id[MAXS], li[MAXS];
		float4 g[MAXS];
		float tout[MAXS];
#pragma unroll
		for (uint32_t k = 0; k < MAXS; k++) {
			id[k] = SLOT_EMPTY; li[k] = LOCAL_OUT; tout[k] = (float) INFINITY;
			if (act && fast && k < S) {
				id[k] = a.cid[base + 64 * k];
				li[k] = a.lidx[base + 64 * k];
				g[k] = a.cgeo[base + 64 * k];
			}
		}
#pragma unroll
		for (uint32_t k = 0; k < MAXS; k++)
			if (act && fast && k < S && li[k] == LOCAL_OUT) tout[k] = od[id[k] & ID30_MASK];
		sd[0][tid] = old;
		__syncthreads();
		uint32_t b = 0;
		for (uint32_t it = 0; it < a.local_iters; it++) {
			const float cur = sd[b][tid];
			float best = cur;
			if (act && fast) {
				float t[MAXS];
#pragma unroll
				for (uint32_t k = 0; k < MAXS; k++) t[k] = (k < S) ? lb_gather(li[k], sd[b], tout[k]) : (float) INFINITY;
#pragma unroll
				for (uint32_t k = 0; k + 1 < MAXS; k++) {
					if (k + 1 < S && !(id[k + 1] & SLOT_EMPTY)) {
						const float p = (id[k] & F64_FLAG)
						                    ? tri_flagged(a, base + 64 * k, t[k], t[k + 1])
						                    : update_stable<float>(g[k].x, g[k + 1].x, g[k].y, g[k].z, g[k].w, t[k],
						                                           t[k + 1], a.h, a.inv_h);
						if (p < best) best = p;
					}
				}
			} else if (act) {
				// Fans longer than MAXS: stream the slots.
				uint32_t ida = a.cid[base];
				uint32_t lia = a.lidx[base];
				float4 ga = a.cgeo[base];
				float ta = lia != LOCAL_OUT ? sd[b][lia] : od[ida & ID30_MASK];
				for (uint32_t k = 1; k < S; k++) {
					const uint32_t idb = a.cid[base + 64 * k];
					if (idb & SLOT_EMPTY) break;
					const uint32_t lib = a.lidx[base + 64 * k];
					const float4 gb = a.cgeo[base + 64 * k];
					const float tb = lib != LOCAL_OUT ? sd[b][lib] : od[idb & ID30_MASK];
					const float p = (ida & F64_FLAG) ? tri_flagged(a, base + 64 * (k - 1), ta, tb)
					                                 : update_stable<float>(ga.x, gb.x, ga.y, ga.z, ga.w, ta, tb, a.h,
					                                                        a.inv_h);
					if (p < best) best = p;
					ida = idb; ga = gb; ta = tb;
				}
			}
			sd[b ^ 1u][tid] = best;
			const int changed = __syncthreads_or(best < cur);
			b ^= 1u;
			if (!changed) break;
		}
		if (act) {
			const float v = sd[b][tid];
			nd[pos] = v;
			if (pos >= cs && pos < ce) bad += !(fabs((double) v - (double) old) / (double) old < a.tol);
		}
		__syncthreads();
	}
	return bad;
}

__device__ __forceinline__ void lb_window(const KArgs &a, const Slot &s, uint32_t &ws, uint32_t &we, uint32_t &cs,
                                          uint32_t &ce)
{
	if (s.phase == 0) {
		ws = a.limits[s.i]; we = a.limits[s.j]; cs = ws; ce = a.limits[s.i + 1];
	} else {
		ws = 0; we = a.nr; cs = 0; ce = a.nr;
	}
}

// Band-schedule sweep k (k4 = k & 3).
__global__ __launch_bounds__(256) void k_lb_sweep(KArgs a, uint32_t k4)
{
	PState *st = a.st;
	const uint32_t kp = (k4 + 3u) & 3u;
	const Slot prev = st->slot[kp];
	const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
	if (prev.done) {
		if (lead) st->slot[k4] = prev;
		return;
	}
	Slot cur = prev;
	cur.fresh = 0;
	if (!prev.fresh) {
		const uint32_t nbad = st->bad[kp];
		uint32_t ws, we, cs, ce;
		lb_window(a, prev, ws, we, cs, ce);
		if (prev.phase == 0) {
			const bool last = prev.j == a.nl - 1;
			const bool retire = nbad == 0 && (prev.i + 1 < prev.j || last);
			cur.i = prev.i + (retire ? 1u : 0u);
			cur.j = prev.j + (last ? 0u : 1u);
			if (!(cur.i < cur.j) || !(prev.iter + 1 < a.max_iter)) { cur.phase = 1; cur.judge = 0; }
		} else {
			cur.judge = prev.judge + 1;
			cur.done = nbad == 0 || cur.judge >= a.judge_cap;
		}
		cur.iter = prev.iter + 1;
		cur.d = prev.d ^ 1u;
		if (lead) {
			if (a.trace) {
				uint32_t *r = a.trace + 8 * (size_t) prev.iter;
				r[0] = prev.iter + 1; r[1] = prev.phase ? 0xFFFFFFFFu : prev.i; r[2] = prev.j; r[3] = ws; r[4] = we;
				r[5] = ce - cs; r[6] = ce - cs - nbad; r[7] = prev.phase;
			}
			st->relax += we - ws;
			st->cond += ce - cs;
			if (a.fcum) st->f64 += a.fcum[we] - a.fcum[ws];
			st->sweeps = cur.iter;
		}
	}
	if (lead) {
		st->slot[k4] = cur;
		st->bad[(k4 + 1u) & 3u] = 0;
		if (cur.done) { st->final_d = cur.d; st->done = 1; }
	}
	if (cur.done) return;
	uint32_t ws, we, cs, ce;
	lb_window(a, cur, ws, we, cs, ce);
	const float *od = (const float *) (cur.d ? a.dist1 : a.dist0);
	float *nd = (float *) (cur.d ? a.dist0 : a.dist1);
	const uint32_t bad = wave_sum(lb_tiles(a, ws, we, cs, ce, od, nd));
	if ((threadIdx.x & 63u) == 0 && bad) atomicAdd(&st->bad[k4], bad);
}

template <typename T>
__global__ void k_init(T *d0, T *d1, uint32_t n, PState *st, uint32_t done)
{
	const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
	for (uint32_t v = t; v < n; v += gridDim.x * blockDim.x) { d0[v] = (T) INFINITY; d1[v] = (T) INFINITY; }
	if (t == 0) {
		uint32_t *sw = (uint32_t *) st;
		for (uint32_t q = 0; q < sizeof(PState) / 4; q++) sw[q] = 0;
		// The state "before" sweep 0 lives in slot 3 and is taken as is.
		st->slot[3].i = 1; st->slot[3].j = 2; st->slot[3].fresh = 1; st->slot[3].done = done;
		st->done = done;
	}
}

template <typename T>
__global__ void k_sources(T *d0, T *d1, const uint32_t *src, uint32_t ns)
{
	const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
	if (t < ns) { d0[src[t]] = (T) 0; d1[src[t]] = (T) 0; }
}

template <typename T>
__global__ void k_to_f64(const T *in, double *out, uint32_t n)
{
	const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
	for (uint32_t v = t; v < n; v += gridDim.x * blockDim.x) out[v] = (double) in[v];
}

}  // namespace gp

// ---------------------------------------------------------------------------
// Solver plan

struct gp_plan {
	gp::Che *mesh = nullptr;                 // borrowed
	std::unique_ptr<gp::Che> coal;           // ring-numbered mesh (coalesce)
	std::vector<uint32_t> inv;               // original -> ring numbering
	gp::DeviceMesh *dm = nullptr;
	gp_ptp_opts opts{};
	uint32_t n = 0, nv_orig = 0, nl = 0, ns = 0, nr = 0, max_iter = 0, grid = 0, block = 256;
	uint32_t nchunks = 0, max_slots = 0, judge_cap = 0;
	unsigned long long n_flagged = 0;
	bool lb = false;
	uint16_t *d_lidx = nullptr;
	uint32_t *d_order = nullptr, *d_src = nullptr, *d_trace = nullptr;
	uint32_t *d_limits = nullptr, *d_coff = nullptr, *d_cid = nullptr;
	unsigned long long *d_fcum = nullptr;
	float4 *d_cgeo = nullptr;
	double4 *d_cgeo64 = nullptr, *d_cx64 = nullptr;
	void *d_dist0 = nullptr, *d_dist1 = nullptr;
	double *d_out = nullptr;
	gp::PState *d_state = nullptr, *h_state = nullptr;
	hipStream_t stream = nullptr;
	hipEvent_t e0 = nullptr, e1 = nullptr;
	hipGraphExec_t g_big = nullptr, g_small = nullptr;
	bool has_run = false;

	~gp_plan()
	{
		for (void *p : { (void *) d_limits, (void *) d_order, (void *) d_src, (void *) d_trace, (void *) d_coff,
		                 (void *) d_cid, (void *) d_cgeo, (void *) d_cgeo64, (void *) d_cx64, d_dist0, d_dist1,
		                 (void *) d_out, (void *) d_state, (void *) d_fcum, (void *) d_lidx })
			(void) hipFree(p);
		if (g_big) (void) hipGraphExecDestroy(g_big);
		if (g_small) (void) hipGraphExecDestroy(g_small);
		if (h_state) (void) hipHostFree(h_state);
		if (e0) (void) hipEventDestroy(e0);
		if (e1) (void) hipEventDestroy(e1);
		if (stream) (void) hipStreamDestroy(stream);
	}
};

namespace gp {

template <typename V>
static int upload(V *&dst, const std::vector<typename std::remove_const<V>::type> &src)
{
	const size_t bytes = sizeof(V) * std::max<size_t>(src.size(), 1);
	if (hipMalloc(&dst, bytes) != hipSuccess) return fail(GP_ERR_NOMEM, "device allocation of %zu bytes failed", bytes);
	if (!src.empty()) GP_HIP(hipMemcpy(dst, src.data(), sizeof(V) * src.size(), hipMemcpyHostToDevice));
	return GP_OK;
}

// Builds a plan from given rings (limits/sorted in the original numbering).
int plan_create_rings(Che *mesh, const uint32_t *src, uint32_t ns, const uint32_t *limits, uint32_t nl,
                      const uint32_t *sorted, const gp_ptp_opts &opts, gp_plan **out)
{
	*out = nullptr;
	if (!mesh || !src || !limits || !sorted) return fail(GP_ERR_INVALID, "plan: null argument");
	if (ns == 0 || nl < 2) return fail(GP_ERR_INVALID, "plan: no sources / empty rings");
	const uint32_t nr = limits[nl - 1];
	if (nr > mesh->nv) return fail(GP_ERR_INVALID, "plan: limits exceed the vertex count");
	// The kernels index with these values: check them all on the host.
	if (limits[0] != 0) return fail(GP_ERR_INVALID, "plan: limits[0] must be 0");
	for (uint32_t q = 1; q < nl; q++)
		if (limits[q] < limits[q - 1]) return fail(GP_ERR_INVALID, "plan: limits must be non-decreasing");
	for (uint32_t q = 0; q < nr; q++)
		if (sorted[q] >= mesh->nv) return fail(GP_ERR_INVALID, "plan: sorted[%u] out of range", q);
	for (uint32_t q = 0; q < ns; q++)
		if (src[q] >= mesh->nv) return fail(GP_ERR_INVALID, "plan: source %u out of range", src[q]);
	auto p = std::make_unique<gp_plan>();
	p->mesh = mesh;
	p->opts = opts;
	p->nv_orig = mesh->nv;
	p->nl = nl;
	p->ns = ns;
	p->nr = nr;
	std::vector<uint32_t> srcm(src, src + ns);
	Che *use = mesh;
	if (opts.coalesce) {
		p->coal = std::make_unique<Che>();
		int rc = coalesce(*mesh, sorted, nr, *p->coal, p->inv);
		if (rc) return rc;
		for (auto &s : srcm) {
			if (p->inv[s] == NIL) return fail(GP_ERR_INVALID, "plan: source not in ring 0");
			s = p->inv[s];
		}
		use = p->coal.get();
	}
	int rc = upload_mesh(*use);
	if (rc) return rc;
	p->dm = use->dev;
	p->n = use->nv;
	const bool f64 = opts.precision == GP_PREC_F64;
	p->lb = opts.schedule != GP_SCHED_RINGS;
	if (p->lb && (f64 || !opts.coalesce)) return fail(GP_ERR_INVALID, "band schedules run in FP32 on a coalesced mesh");
	if (p->lb && (opts.local_iters == 0 || opts.local_iters > 1024)) return fail(GP_ERR_INVALID, "local_iters out of range");
	Layout L;
	rc = build_layout(*use, opts.coalesce ? nullptr : sorted, nr, f64, L, p->lb ? TILE : 0);
	if (rc) return rc;
	p->nchunks = L.nchunks;
	p->max_slots = L.max_slots;
	p->n_flagged = L.n_flagged;
	p->judge_cap = opts.judge_cap ? opts.judge_cap : 64;
	p->max_iter = opts.max_sweeps ? opts.max_sweeps : (p->lb ? 4 * nl + 64 : (nl << 1));
	p->block = 256;
	// One wave per chunk, at most as many workgroups as fit on the device at
	// once (idle workgroups of small windows then cost no second dispatch round).
	int dev = 0, ncu = 0, occ = 0;
	GP_HIP(hipGetDevice(&dev));
	GP_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
	if (p->lb) GP_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_lb_sweep, 256, 0));
	else if (f64) GP_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_sweep<true>, 256, 0));
	else GP_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_sweep<false>, 256, 0));
	const uint32_t gmax = (uint32_t) ncu * (uint32_t) std::max(1, occ);
	p->grid = opts.grid ? opts.grid : std::max<uint32_t>(1, std::min<uint32_t>(gmax, (L.nchunks + 3) / 4));
	if (p->lb) {
		if ((rc = upload(p->d_lidx, L.lidx))) return rc;
	}

	const size_t es = f64 ? sizeof(double) : sizeof(float);
	GP_HIP(hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking));
	GP_HIP(hipEventCreate(&p->e0));
	GP_HIP(hipEventCreate(&p->e1));
	std::vector<uint32_t> lim(limits, limits + nl);
	if ((rc = upload(p->d_limits, lim)) || (rc = upload(p->d_src, srcm)) || (rc = upload(p->d_coff, L.coff)) ||
	    (rc = upload(p->d_cid, L.cid)))
		return rc;
	auto as_d4 = [](const std::vector<Geo64> &v) {
		std::vector<double4> o(v.size());
		std::memcpy(o.data(), v.data(), sizeof(double4) * v.size());
		return o;
	};
	if (f64) {
		if ((rc = upload(p->d_cx64, as_d4(L.cx64)))) return rc;
	} else {
		std::vector<float4> cg(L.cgeo.size());
		std::memcpy(cg.data(), L.cgeo.data(), sizeof(float4) * cg.size());
		if ((rc = upload(p->d_cgeo, cg)) || (rc = upload(p->d_cgeo64, as_d4(L.cgeo64))) ||
		    (rc = upload(p->d_fcum, L.fcum)))
			return rc;
	}
	if (!opts.coalesce) {
		std::vector<uint32_t> ord(sorted, sorted + nr);
		if ((rc = upload(p->d_order, ord))) return rc;
	}
	if (hipMalloc(&p->d_dist0, es * p->n) != hipSuccess || hipMalloc(&p->d_dist1, es * p->n) != hipSuccess ||
	    hipMalloc(&p->d_out, sizeof(double) * p->n) != hipSuccess ||
	    hipMalloc(&p->d_state, sizeof(PState)) != hipSuccess ||
	    hipMalloc(&p->d_trace, sizeof(uint32_t) * 8 * (size_t) (p->max_iter + p->judge_cap + 1)) != hipSuccess ||
	    hipHostMalloc(&p->h_state, sizeof(PState)) != hipSuccess)
		return fail(GP_ERR_NOMEM, "plan: device allocation failed");
	*out = p.release();
	return GP_OK;
}

static KArgs make_args(const gp_plan *p)
{
	KArgs a{};
	a.limits = p->d_limits; a.coff = p->d_coff; a.cid = p->d_cid; a.cgeo = p->d_cgeo; a.cgeo64 = p->d_cgeo64;
	a.cx64 = p->d_cx64; a.fcum = p->d_fcum; a.order = p->d_order; a.dist0 = p->d_dist0; a.dist1 = p->d_dist1;
	a.st = p->d_state; a.trace = p->d_trace;
	a.tol = p->opts.tol; a.h = p->dm->h; a.inv_h = p->dm->inv_h;
	a.nl = p->nl; a.max_iter = p->max_iter;
	a.lidx = p->d_lidx; a.nr = p->nr; a.local_iters = p->opts.local_iters; a.judge_cap = p->judge_cap;
	return a;
}

// Enqueues sweeps k0 .. k0+count-1 (their k & 3 slots).
static void enqueue_sweeps(const gp_plan *p, const KArgs &a, uint32_t k0, uint32_t count)
{
	for (uint32_t q = 0; q < count; q++) {
		const uint32_t k4 = (k0 + q) & 3u;
		if (p->lb) k_lb_sweep<<<p->grid, p->block, 0, p->stream>>>(a, k4);
		else if (p->opts.precision == GP_PREC_F64) k_sweep<true><<<p->grid, p->block, 0, p->stream>>>(a, k4);
		else k_sweep<false><<<p->grid, p->block, 0, p->stream>>>(a, k4);
	}
}

static int capture(gp_plan *p, const KArgs &a, uint32_t count, hipGraphExec_t *out)
{
	hipGraph_t g = nullptr;
	GP_HIP(hipStreamBeginCapture(p->stream, hipStreamCaptureModeThreadLocal));
	enqueue_sweeps(p, a, 0, count);
	GP_HIP(hipStreamEndCapture(p->stream, &g));
	const hipError_t e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
	(void) hipGraphDestroy(g);
	if (e != hipSuccess) return fail(GP_ERR_HIP, "hipGraphInstantiate failed: %s", hipGetErrorString(e));
	return GP_OK;
}

int plan_run(gp_plan *p, gp_ptp_stats *stats)
{
	if (!p) return fail(GP_ERR_INVALID, "plan is null");
	const auto t0 = std::chrono::steady_clock::now();
	const bool f64 = p->opts.precision == GP_PREC_F64;
	const uint32_t done0 = p->nl < 3 ? 1u : 0u;
	const uint32_t ig = std::max<uint32_t>(1, std::min<uint32_t>(1024, (p->n + 255) / 256));
	const KArgs a = make_args(p);
	constexpr uint32_t BIG = 64, SMALL = 8;
	const bool graph = p->opts.launch_mode == GP_LAUNCH_GRAPH;
	if (graph && !p->g_big) {
		int rc = capture(p, a, BIG, &p->g_big);
		if (!rc) rc = capture(p, a, SMALL, &p->g_small);
		if (rc) return rc;
	}
	if (f64) {
		k_init<double><<<ig, 256, 0, p->stream>>>((double *) p->d_dist0, (double *) p->d_dist1, p->n, p->d_state, done0);
		k_sources<double><<<(p->ns + 255) / 256, 256, 0, p->stream>>>((double *) p->d_dist0, (double *) p->d_dist1, p->d_src, p->ns);
	} else {
		k_init<float><<<ig, 256, 0, p->stream>>>((float *) p->d_dist0, (float *) p->d_dist1, p->n, p->d_state, done0);
		k_sources<float><<<(p->ns + 255) / 256, 256, 0, p->stream>>>((float *) p->d_dist0, (float *) p->d_dist1, p->d_src, p->ns);
	}
	GP_HIP(hipGetLastError());
	GP_HIP(hipEventRecord(p->e0, p->stream));
	// Replays until the device reports done.  One sweep more than the sweep
	// count runs (it records the last sweep's outcome).  j grows by one ring
	// per sweep until |limits|-1, so at least nl-1-j sweeps remain: big graphs
	// while that bound allows, small ones near the end.
	uint32_t launched = 0, j_seen = 2;
	const uint32_t cap = p->max_iter + (p->lb ? p->judge_cap : 0);
	while (!done0 && launched <= cap) {
		const bool big = p->nl - 1 > j_seen + BIG;
		const uint32_t cnt = big ? BIG : SMALL;
		if (graph) GP_HIP(hipGraphLaunch(big ? p->g_big : p->g_small, p->stream));
		else enqueue_sweeps(p, a, launched, cnt);
		launched += cnt;
		GP_HIP(hipGetLastError());
		GP_HIP(hipMemcpyAsync(p->h_state, p->d_state, sizeof(PState), hipMemcpyDeviceToHost, p->stream));
		GP_HIP(hipStreamSynchronize(p->stream));
		if (p->h_state->done) break;
		j_seen = p->h_state->slot[(launched - 1) & 3u].j;
	}
	if (!done0 && !p->h_state->done) return fail(GP_ERR_KERNEL, "ptp: schedule did not finish in %u sweeps", launched);
	GP_HIP(hipEventRecord(p->e1, p->stream));
	GP_HIP(hipMemcpyAsync(p->h_state, p->d_state, sizeof(PState), hipMemcpyDeviceToHost, p->stream));
	GP_HIP(hipStreamSynchronize(p->stream));
	float ms = 0;
	GP_HIP(hipEventElapsedTime(&ms, p->e0, p->e1));
	p->has_run = true;
	if (stats) {
		stats->sweeps = p->h_state->sweeps;
		stats->n_limits = p->nl;
		stats->relaxations = p->h_state->relax;
		stats->cond_checks = p->h_state->cond;
		stats->f64_updates = p->h_state->f64;
		stats->kernel_ms = ms;
		stats->total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
	}
	return GP_OK;
}

int plan_result(gp_plan *p, double *dist)
{
	if (!p || !dist) return fail(GP_ERR_INVALID, "plan_result: null argument");
	if (!p->has_run) return fail(GP_ERR_INVALID, "plan_result: plan has not run");
	const uint32_t d = p->h_state->final_d;
	const bool last_read = p->opts.result_buffer == GP_RESULT_LAST_READ;
	const void *buf = (last_read ? !d : d) ? p->d_dist1 : p->d_dist0;
	const uint32_t g = std::max<uint32_t>(1, std::min<uint32_t>(1024, (p->n + 255) / 256));
	if (p->opts.precision == GP_PREC_F64)
		k_to_f64<double><<<g, 256, 0, p->stream>>>((const double *) buf, p->d_out, p->n);
	else
		k_to_f64<float><<<g, 256, 0, p->stream>>>((const float *) buf, p->d_out, p->n);
	GP_HIP(hipGetLastError());
	if (p->coal) {
		std::vector<double> tmp(p->n);
		GP_HIP(hipMemcpyAsync(tmp.data(), p->d_out, sizeof(double) * p->n, hipMemcpyDeviceToHost, p->stream));
		GP_HIP(hipStreamSynchronize(p->stream));
		for (uint32_t v = 0; v < p->nv_orig; v++) dist[v] = p->inv[v] != NIL ? tmp[p->inv[v]] : INFINITY;
	} else {
		GP_HIP(hipMemcpyAsync(dist, p->d_out, sizeof(double) * p->n, hipMemcpyDeviceToHost, p->stream));
		GP_HIP(hipStreamSynchronize(p->stream));
	}
	return GP_OK;
}

int plan_trace(gp_plan *p, uint32_t *out, uint32_t cap, uint32_t *n)
{
	if (!p || !p->has_run) return fail(GP_ERR_INVALID, "plan_trace: plan has not run");
	const uint32_t s = p->h_state->sweeps;
	if (n) *n = s;
	const uint32_t c = std::min(s, cap);
	if (c && out) {
		GP_HIP(hipMemcpyAsync(out, p->d_trace, sizeof(uint32_t) * 8 * (size_t) c, hipMemcpyDeviceToHost, p->stream));
		GP_HIP(hipStreamSynchronize(p->stream));
	}
	return GP_OK;
}

}  // namespace gp

extern "C" void gp_plan_destroy(gp_plan *p) { delete p; }
