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
#include "levels.hpp"
#include "sweep_dev.hpp"
#include "update_step.hpp"

namespace gp {

// Vertex-major slot tables of a mesh in its own numbering (slot k of vertex v
// at fan_off[v] + k; a closed fan stored without its repeated first entry
// leaves its last fan entry unused), the source that device-built plans and
// farthest-point sampling gather their ring-ordered tables from.  Built on the
// device once per mesh and precision.
struct MasterLayout {
	uint32_t *cid = nullptr;
	float4 *cgeo = nullptr;
	double4 *cgeo64 = nullptr;
	Edge64 *cx64 = nullptr;
	uint32_t max_slots = 0;
	VertexTables view(const DeviceMesh *dm) const
	{
		return VertexTables{ dm->fan_off, dm->closed, cid, cgeo, cgeo64, cx64 };
	}
	~MasterLayout()
	{
		for (void *p : { (void *) cid, (void *) cgeo, (void *) cgeo64, (void *) cx64 }) (void) hipFree(p);
	}
};

void device_mesh_free(DeviceMesh *d)
{
	if (!d) return;
	(void) hipFree(d->fan_off);
	(void) hipFree(d->fan);
	(void) hipFree(d->pos);
	(void) hipFree(d->closed);
	delete d->master[0];
	delete d->master[1];
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
//
// A slot also carries its sweep's window in positions (ws, wc, we: start, end
// of the convergence-tested ring, end) and, written at the end of the sweep,
// the two windows the next sweep can have (ring i kept or retired: cs/cc[0|1],
// end ce), so that the next sweep knows its window from the slot alone,
// without a dependent read of limits.
struct Slot {
	uint32_t i, j, d, iter, done, fresh, ws, wc;
	uint32_t we, cs0, cc0, cs1, cc1, ce, pad0, pad1;
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
	uint32_t sb;               // > 0: every chunk has sb slots (chunk c at row c * sb; coff unused)
	const uint32_t *cid;
	const float4 *cgeo;
	const double4 *cgeo64;
	const Edge64 *cx64;
	const unsigned long long *fcum;  // FP64-update triangles of positions [0, pos)
	const uint32_t *order;     // position -> vertex; null when the mesh is numbered by rings
	void *dist0, *dist1;
	PState *st;
	uint32_t *trace;           // 8 words per sweep, or null
	double tol;
	float h, inv_h;
	uint32_t nl, max_iter;
	uint32_t k2;               // 1: K/2 limit (the drivers); 0: the uncapped iter_error loop
	uint32_t *cl0, *cl1;       // cluster labels (double-buffered like the distances), or null
};

// Relaxes the window [start, end) (positions); returns the wave's count of
// non-converged vertices among positions [start, cend).
template <bool F64, bool CL, uint32_t MS, bool G64>
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
		const uint32_t cb = a.sb ? c * a.sb : a.coff[c], S = a.sb ? a.sb : a.coff[c + 1] - cb;
		bool lbad = false;
		if (pos >= start && pos < end) {
			const uint32_t v = a.order ? a.order[pos] : pos;
			const T old = od[v];
			const size_t base = (size_t) cb * 64 + lane;
			uint32_t bk = NO_TRI;
			const T best = relax_lane<F64, 64, true, CL, MS, G64>(a, base, S, od, old, &bk);
			nd[v] = best;
			if constexpr (CL) {
				// The label of the nearer of the winning triangle's two other
				// vertices (old values; ties to the first fan entry), as the
				// reference's GPU relax does (src/cuda/geodesics_ptp.cu:266-276).
				const uint32_t *ocl = od == (const T *) a.dist0 ? a.cl0 : a.cl1;
				uint32_t *ncl = od == (const T *) a.dist0 ? a.cl1 : a.cl0;
				uint32_t c = ocl[v];
				if (bk != NO_TRI) {
					const uint32_t ia = a.cid[base + 64 * (size_t) bk] & SLOT_ID_MASK;
					const uint32_t nb = bk + 1 < S ? a.cid[base + 64 * (size_t) (bk + 1)] : SLOT_EMPTY;
					const uint32_t ib = ((nb & SLOT_EMPTY) ? a.cid[base] : nb) & SLOT_ID_MASK;
					c = od[ib] < od[ia] ? ocl[ib] : ocl[ia];
				}
				ncl[v] = c;
			}
			if (pos < cend) lbad = !(fabs((double) best - (double) old) / (double) old < a.tol);
		}
		bad += (uint32_t) __popcll(__ballot(lbad));
	}
	return bad;
}

// One wave per workgroup: waves of a sweep finish and free their slots
// independently, and the XCD split deals the window out in finer pieces
// (measured on the 1M field against 4 waves per workgroup: 2.5 % faster on
// the reference schedule, 4 % on the learned bands).
constexpr uint32_t SWEEP_BLOCK = 64;

// Sweep k (k4 = k & 3) of a run: advance the schedule, then relax its window.
template <bool F64, bool CL, uint32_t MS, bool G64 = false>
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
		const uint32_t pi = a.k2 && prev.i < (prev.j >> 1) ? (prev.j >> 1) : prev.i;
		const bool adv = nbad == 0;
		cur.i = pi + (adv ? 1u : 0u);
		cur.j = prev.j + (prev.j < a.nl - 1 ? 1u : 0u);
		cur.d = prev.d ^ 1u;
		cur.iter = prev.iter + 1;
		cur.done = !(cur.i < cur.j) || !(cur.iter < a.max_iter);
		cur.ws = adv ? prev.cs1 : prev.cs0;
		cur.wc = adv ? prev.cc1 : prev.cc0;
		cur.we = prev.ce;
		if (lead) {
			const uint32_t l0 = prev.ws, lj = prev.we;
			const uint32_t ncond = prev.wc - prev.ws;
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
	const uint32_t i = a.k2 && cur.i < (cur.j >> 1) ? (cur.j >> 1) : cur.i;   // K/2 limit band size
	if (prev.fresh && !cur.done) {
		cur.ws = a.limits[i]; cur.wc = a.limits[i + 1]; cur.we = a.limits[cur.j];
	}
	// The next sweep's two possible windows (read here, stored at the end).
	uint32_t nx[5] = { 0, 0, 0, 0, 0 };
	if (lead && !cur.done) {
		const uint32_t jn = cur.j + (cur.j < a.nl - 1 ? 1u : 0u);
		for (uint32_t adv = 0; adv < 2; adv++) {
			const uint32_t in = i + adv;
			const uint32_t iw = a.k2 && in < (jn >> 1) ? (jn >> 1) : in;
			if (in < jn) { nx[2 * adv] = a.limits[iw]; nx[2 * adv + 1] = a.limits[iw + 1]; }
		}
		nx[4] = a.limits[jn];
	}
	if (lead) {
		st->slot[k4] = cur;
		st->bad[(k4 + 1u) & 3u] = 0;
		if (cur.done) { st->final_d = cur.d; st->done = 1; }
	}
	if (cur.done) return;
	const uint32_t start = cur.ws, cend = cur.wc, end = cur.we;
	const T *od = (const T *) (cur.d ? a.dist1 : a.dist0);
	T *nd = (T *) (cur.d ? a.dist0 : a.dist1);
	const uint32_t bad = sweep_body<F64, CL, MS, G64>(a, start, end, cend, od, nd);
	if ((threadIdx.x & 63u) == 0 && bad) atomicAdd(&st->bad[k4], bad);
	if (lead) {
		Slot *sl = &st->slot[k4];
		sl->cs0 = nx[0]; sl->cc0 = nx[1]; sl->cs1 = nx[2]; sl->cc1 = nx[3]; sl->ce = nx[4];
	}
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

// Cluster labels: NIL everywhere, then source q gets q + 1
// (src/cuda/geodesics_ptp.cu:186-195; the reference leaves the others unset).
__global__ void k_cl_init(uint32_t *c0, uint32_t *c1, uint32_t n)
{
	for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) c0[v] = c1[v] = GP_NIL;
}
__global__ void k_cl_sources(uint32_t *c0, uint32_t *c1, const uint32_t *src, uint32_t ns)
{
	const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
	if (t < ns) c0[src[t]] = c1[src[t]] = t + 1;
}

template <typename T>
__global__ void k_to_f64(const T *in, double *out, uint32_t n)
{
	const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
	for (uint32_t v = t; v < n; v += gridDim.x * blockDim.x) out[v] = (double) in[v];
}

// Device-built plans: the result in the caller's numbering.
__global__ void k_from_positions(const double *in, const uint32_t *inv, uint32_t nv, double *out)
{
	for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < nv; v += gridDim.x * blockDim.x)
		out[v] = inv[v] != GP_NIL ? in[inv[v]] : (double) INFINITY;
}

__global__ void k_fill_u32(uint32_t *p, uint32_t n, uint32_t x)
{
	for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) p[v] = x;
}

// inv[sorted[p]] = p for the nr reached vertices; src_pos = 0 .. ns - 1 (ring 0
// holds exactly the sources).
// Band plans built on the device: a reached vertex other than a source (BFS
// level 0) takes band 1 + floor(dhat / width) (clamped to nv; the host band
// builder's index, che.cpp band_order), the sources band 0.  *bmax gets the
// largest band.
__global__ void k_band_levels(uint32_t *level, const double *dhat, uint32_t nv, double width, uint32_t *bmax)
{
	uint32_t m = 0;
	for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < nv; v += gridDim.x * blockDim.x) {
		const uint32_t l = level[v];
		if (l == GP_NIL || l == 0) continue;
		const double d = dhat[v];
		const double bq = isfinite(d) && d > 0 ? floor(d / width) : 0.0;
		const unsigned long long band = 1ull + (unsigned long long) fmin(bq, 4.0e9);
		const uint32_t b = band < nv ? (uint32_t) band : nv;
		level[v] = b;
		m = max(m, b);
	}
	for (uint32_t o = 32; o > 0; o >>= 1) m = max(m, (uint32_t) __shfl_xor(m, o));
	if ((threadIdx.x & 63u) == 0) atomicMax(bmax, m);
}

// Drops the empty bands from limits[0 .. *nl - 1] (a level with no vertex
// repeats its start), as the host band builder never opens one; tmp holds
// *nl words.  One workgroup of 1024 threads.
__global__ __launch_bounds__(1024) void k_compact_limits(uint32_t *limits, uint32_t *nl, uint32_t *tmp)
{
	__shared__ uint32_t part[1024];
	const uint32_t n = *nl;
	uint32_t carry = 0;
	for (uint32_t b0 = 0; b0 < n; b0 += 1024) {
		const uint32_t q = b0 + threadIdx.x;
		const uint32_t keep = (q < n && (q == 0 || limits[q] != limits[q - 1])) ? 1u : 0u;
		part[threadIdx.x] = keep;
		__syncthreads();
		for (uint32_t o = 1; o < 1024; o <<= 1) {
			const uint32_t y = threadIdx.x >= o ? part[threadIdx.x - o] : 0u;
			__syncthreads();
			part[threadIdx.x] += y;
			__syncthreads();
		}
		if (keep) tmp[carry + part[threadIdx.x] - 1] = limits[q];
		const uint32_t tot = part[1023];
		__syncthreads();
		carry += tot;
	}
	__syncthreads();
	for (uint32_t q = threadIdx.x; q < carry; q += 1024) limits[q] = tmp[q];
	if (threadIdx.x == 0) *nl = carry;
}

__global__ void k_positions(const uint32_t *sorted, uint32_t nr, uint32_t *inv, uint32_t *src_pos, uint32_t ns)
{
	for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < nr; q += gridDim.x * blockDim.x) {
		inv[sorted[q]] = q;
		if (q < ns) src_pos[q] = q;
	}
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
	uint32_t nchunks = 0, max_slots = 0, sb = 0;
	bool ms8 = false;                        // every fan fits 8 slots: the 8-slot kernels
	bool g64 = false;                        // FP32 sweeps read the FP64 geometry table (GP_GEO_FP64)
	unsigned long long n_flagged = 0;
	gp::PatchPlan *pp = nullptr;             // patch / band schedules
	uint32_t *d_order = nullptr, *d_src = nullptr, *d_trace = nullptr;
	uint32_t *d_limits = nullptr, *d_coff = nullptr, *d_cid = nullptr;
	unsigned long long *d_fcum = nullptr;
	float4 *d_cgeo = nullptr;
	double4 *d_cgeo64 = nullptr;
	gp::Edge64 *d_cx64 = nullptr;
	void *d_dist0 = nullptr, *d_dist1 = nullptr;
	uint32_t *d_cl0 = nullptr, *d_cl1 = nullptr;   // cluster labels (opts.clusters)
	uint32_t *d_inv = nullptr;                // device-built plans: original vertex -> position (NIL: unreached)
	double *d_out = nullptr;
	gp::PState *d_state = nullptr, *h_state = nullptr;   // h_state: 2 pinned snapshots
	hipStream_t stream = nullptr;
	hipEvent_t e0 = nullptr, e1 = nullptr, ev[2] = { nullptr, nullptr };
	hipGraphExec_t g_big = nullptr, g_small = nullptr;
	bool has_run = false;

	~gp_plan()
	{
		for (void *p : { (void *) d_limits, (void *) d_order, (void *) d_src, (void *) d_trace, (void *) d_coff,
		                 (void *) d_cid, (void *) d_cgeo, (void *) d_cgeo64, (void *) d_cx64, d_dist0, d_dist1,
		                 (void *) d_out, (void *) d_state, (void *) d_fcum, (void *) d_cl0, (void *) d_cl1,
		                 (void *) d_inv })
			(void) hipFree(p);
		if (pp) gp::patch_destroy(pp);
		if (g_big) (void) hipGraphExecDestroy(g_big);
		if (g_small) (void) hipGraphExecDestroy(g_small);
		if (h_state) (void) hipHostFree(h_state);
		if (e0) (void) hipEventDestroy(e0);
		if (e1) (void) hipEventDestroy(e1);
		for (hipEvent_t e : ev)
			if (e) (void) hipEventDestroy(e);
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

// The sweep kernel's launch width: one wave per chunk, at most as many
// workgroups as fit on the device at once (idle workgroups of small windows
// then cost no second dispatch round).  Fans of up to 8 slots (interior
// valence up to 8 once closed fans drop their repeat) run a kernel with 8
// register-resident slots.
static int sweep_grid(gp_plan *p, bool f64)
{
	int dev = 0, ncu = 0, occ = 0;
	GP_HIP(hipGetDevice(&dev));
	GP_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
	p->ms8 = p->max_slots <= 8;
	const void *kfn = f64 ? (p->ms8 ? (const void *) k_sweep<true, false, 8> : (const void *) k_sweep<true, false, MAXS>)
	                : p->g64 ? (p->ms8 ? (const void *) k_sweep<false, false, 8, true>
	                                   : (const void *) k_sweep<false, false, MAXS, true>)
	                         : (p->ms8 ? (const void *) k_sweep<false, false, 8> : (const void *) k_sweep<false, false, MAXS>);
	GP_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kfn, (int) SWEEP_BLOCK, 0));
	const uint32_t gmax = (uint32_t) ncu * (uint32_t) std::max(1, occ);
	p->grid = p->opts.grid ? p->opts.grid
	                       : std::max<uint32_t>(1, std::min<uint32_t>(gmax, (p->nchunks * 64 + SWEEP_BLOCK - 1) / SWEEP_BLOCK));
	return GP_OK;
}

static int plan_stream(gp_plan *p)
{
	GP_HIP(hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking));
	GP_HIP(hipEventCreate(&p->e0));
	GP_HIP(hipEventCreate(&p->e1));
	GP_HIP(hipEventCreate(&p->ev[0]));
	GP_HIP(hipEventCreate(&p->ev[1]));
	return GP_OK;
}

// Distances, labels, result, schedule state and trace of a ring plan.
static int plan_run_buffers(gp_plan *p, bool f64)
{
	const size_t es = f64 ? sizeof(double) : sizeof(float);
	if (p->opts.clusters &&
	    (hipMalloc(&p->d_cl0, 4ull * p->n) != hipSuccess || hipMalloc(&p->d_cl1, 4ull * p->n) != hipSuccess))
		return fail(GP_ERR_NOMEM, "plan: cluster allocation failed");
	if (hipMalloc(&p->d_dist0, es * p->n) != hipSuccess || hipMalloc(&p->d_dist1, es * p->n) != hipSuccess ||
	    hipMalloc(&p->d_out, sizeof(double) * p->n) != hipSuccess ||
	    hipMalloc(&p->d_state, sizeof(PState)) != hipSuccess ||
	    hipMalloc(&p->d_trace, sizeof(uint32_t) * 8 * (size_t) (p->max_iter + 1)) != hipSuccess ||
	    hipHostMalloc(&p->h_state, 2 * sizeof(PState)) != hipSuccess)
		return fail(GP_ERR_NOMEM, "plan: device allocation failed");
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
	// Inside each ring the vertices are ordered by slot count (longest fans
	// first, the given order otherwise kept), so that a chunk of 64 positions
	// holds fans of one length and its slot count is its own fans' (a sweep
	// relaxes whole rings from the previous buffer: the order inside a ring
	// changes no value).  Measured on the 1M height field: 10 % fewer bytes
	// and time on the large windows of the band and uncapped schedules.
	std::vector<uint32_t> by_len(sorted, sorted + nr);
	{
		auto slots = [&](uint32_t v) {
			const uint32_t n = mesh->fan_off[v + 1] - mesh->fan_off[v];
			return (n && mesh->fan_closed[v]) ? n - 1 : n;
		};
		for (uint32_t r = 0; r + 1 < nl; r++)
			std::stable_sort(by_len.begin() + limits[r], by_len.begin() + limits[r + 1],
			                 [&](uint32_t x, uint32_t y) { return slots(x) > slots(y); });
		sorted = by_len.data();
	}
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
	if (opts.schedule != GP_SCHED_RINGS) return fail(GP_ERR_INVALID, "plan_create_rings: ring schedule only");
	Layout L;
	rc = build_layout(*use, opts.coalesce ? nullptr : sorted, nr, f64, L, 0, 64, 0, true);
	if (rc) return rc;
	p->nchunks = L.nchunks;
	p->max_slots = L.max_slots;
	p->sb = L.uniform;
	p->n_flagged = L.n_flagged;
	p->g64 = !f64 && (opts.fp32_geometry == GP_GEO_FP64 ||
	                  (opts.fp32_geometry == GP_GEO_AUTO && 4 * L.n_flagged > L.n_slots));
	// The drivers cap the loop at 2*|limits| sweeps (src/geodesics_ptp.cpp:66-67);
	// the uncapped loop (src/cuda/test_geodesics_ptp_coalescence.cu:223-248)
	// has no cap, so a generous one stands in as a guard.
	p->max_iter = opts.max_sweeps ? opts.max_sweeps : (opts.uncapped ? 64 * nl + 64 : (nl << 1));
	p->block = SWEEP_BLOCK;
	if ((rc = sweep_grid(p.get(), f64))) return rc;
	if ((rc = plan_stream(p.get()))) return rc;
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
		if ((rc = upload(p->d_cx64, L.cx64))) return rc;
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
	if ((rc = plan_run_buffers(p.get(), f64))) return rc;
	*out = p.release();
	return GP_OK;
}

// The mesh's vertex-major slot tables, one thread per vertex: the values of
// the host layout builder (che.cpp, build_layout) -- the same FP64 operations
// in the same order, contraction off, so the same bits and the same FP64
// flags.  FP32 solves: neighbour id (F64_FLAG for a flagged corner, SLOT_WRAP
// on slot 0 of a closed fan), FP32 and FP64 triangle geometry in units of the
// mean edge; FP64 solves: neighbour id and the FP64 edge vector.
template <bool F64>
__global__ void k_master_tables(const uint32_t *fan_off, const uint32_t *fan, const uint8_t *closed, const double *gt,
                                uint32_t nv, double ih2, uint32_t *cid, float4 *cgeo, double4 *cgeo64, Edge64 *cx64)
{
#pragma clang fp contract(off)
	for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < nv; v += gridDim.x * blockDim.x) {
		const uint32_t s0 = fan_off[v], n = fan_off[v + 1] - s0;
		const uint32_t ns = (n && closed[v]) ? n - 1 : n;
		const double pv[3] = { gt[3 * (size_t) v], gt[3 * (size_t) v + 1], gt[3 * (size_t) v + 2] };
		for (uint32_t k = 0; k < ns; k++) {
			const size_t at = (size_t) s0 + k;
			const uint32_t a = fan[s0 + k];
			const double *pa = gt + 3 * (size_t) a;
			const double X0[3] = { pa[0] - pv[0], pa[1] - pv[1], pa[2] - pv[2] };
			uint32_t id = a;
			if constexpr (F64) {
				cx64[at] = Edge64{ X0[0], X0[1], X0[2] };
			} else {
				const double q0 = X0[0] * X0[0] + X0[1] * X0[1] + X0[2] * X0[2];
				double4 gd = make_double4(q0 * ih2, 0, 0, 0);
				if (k + 1 < n) {
					const double *pb = gt + 3 * (size_t) fan[s0 + k + 1];
					const double X1[3] = { pb[0] - pv[0], pb[1] - pv[1], pb[2] - pv[2] };
					const double q1 = X1[0] * X1[0] + X1[1] * X1[1] + X1[2] * X1[2];
					const double g = X0[0] * X1[0] + X0[1] * X1[1] + X0[2] * X1[2];
					const double cx = X0[1] * X1[2] - X0[2] * X1[1];
					const double cy = X0[2] * X1[0] - X0[0] * X1[2];
					const double cz = X0[0] * X1[1] - X0[1] * X1[0];
					const double det = cx * cx + cy * cy + cz * cz;
					const double E[3] = { pb[0] - pa[0], pb[1] - pa[1], pb[2] - pa[2] };
					gd.y = g * ih2;
					gd.z = det * ih2 * ih2;
					gd.w = (E[0] * E[0] + E[1] * E[1] + E[2] * E[2]) * ih2;
					const double qq = q0 * q1;
					if (!(qq > 0) || det < FLAG_SIN2_MIN * qq || g < FLAG_COS_MIN * sqrt(qq)) id |= F64_FLAG;
				}
				cgeo[at] = make_float4((float) gd.x, (float) gd.y, (float) gd.z, (float) gd.w);
				cgeo64[at] = gd;
			}
			if (k == 0 && ns < n) id |= SLOT_WRAP;
			cid[at] = id;
		}
	}
}

// Vertex-major slot tables of the mesh (built on the device once per mesh and
// precision, kept with the device mesh) and the longest fan's slot count.
int master_tables(Che &mesh, bool f64, VertexTables &out, uint32_t &max_slots)
{
	int rc = upload_mesh(mesh);
	if (rc) return rc;
	DeviceMesh *dm = mesh.dev;
	MasterLayout *ml = dm->master[f64 ? 1 : 0];
	if (!ml) {
		if (mesh.nv >= SLOT_WRAP) return fail(GP_ERR_CAPACITY, "layout: vertex ids must stay below 2^29");
		auto m = std::make_unique<MasterLayout>();
		for (uint32_t v = 0; v < mesh.nv; v++) {
			const uint32_t n = mesh.fan_off[v + 1] - mesh.fan_off[v];
			m->max_slots = std::max(m->max_slots, (n && mesh.fan_closed[v]) ? n - 1 : n);
		}
		const size_t nf = std::max<size_t>(dm->nfan, 1);
		if (hipMalloc(&m->cid, 4 * nf) != hipSuccess ||
		    (f64 ? hipMalloc(&m->cx64, sizeof(Edge64) * nf)
		         : (hipMalloc(&m->cgeo, sizeof(float4) * nf) == hipSuccess ? hipMalloc(&m->cgeo64, sizeof(double4) * nf)
		                                                                  : hipErrorOutOfMemory)) != hipSuccess)
			return fail(GP_ERR_NOMEM, "layout: device allocation failed");
		const double ih = 1.0 / mesh.mean_edge, ih2 = ih * ih;
		const uint32_t g = std::max<uint32_t>(1, std::min<uint32_t>(2048, (mesh.nv + 255) / 256));
		if (f64)
			k_master_tables<true><<<g, 256>>>(dm->fan_off, dm->fan, dm->closed, dm->pos, mesh.nv, ih2, m->cid, nullptr,
			                                  nullptr, m->cx64);
		else
			k_master_tables<false><<<g, 256>>>(dm->fan_off, dm->fan, dm->closed, dm->pos, mesh.nv, ih2, m->cid,
			                                   m->cgeo, m->cgeo64, nullptr);
		GP_HIP(hipGetLastError());
		GP_HIP(hipDeviceSynchronize());
		ml = dm->master[f64 ? 1 : 0] = m.release();
	}
	out = ml->view(dm);
	max_slots = ml->max_slots;
	return GP_OK;
}

// Plan of the ring schedule on a ring-numbered layout built on the device:
// rings by the frontier BFS and counting sort (no host BFS, no host
// renumbering), tables gathered from the mesh's vertex-major tables
// (k_gather_tables; each ring ordered longest fans first, each chunk as many
// slot rows as its own longest fan).  The rings are the reference's as sets,
// the order inside a ring is not (nothing depends on it); no FP64-update
// count.  GP_ERR_CAPACITY when a fan needs
// more register-held slots than the sweep kernels hold (the caller then takes
// the host-built path).
int plan_create_device(Che *mesh, const uint32_t *src, uint32_t ns, const gp_ptp_opts &opts, gp_plan **out,
                       const double *dhat, double width)
{
	*out = nullptr;
	if (!mesh || !src) return fail(GP_ERR_INVALID, "plan: null argument");
	if (ns == 0) return fail(GP_ERR_INVALID, "plan: no sources");
	if (opts.schedule != GP_SCHED_RINGS || !opts.coalesce || opts.clusters)
		return fail(GP_ERR_INVALID, "plan_create_device: ring schedule, coalesced, without labels only");
	{
		std::vector<uint8_t> seen(mesh->nv, 0);
		for (uint32_t q = 0; q < ns; q++) {
			if (src[q] >= mesh->nv) return fail(GP_ERR_INVALID, "plan: source %u out of range", src[q]);
			if (seen[src[q]]) return fail(GP_ERR_DUPLICATE_SOURCE, "plan: source %u listed twice", src[q]);
			seen[src[q]] = 1;
		}
	}
	int rc = upload_mesh(*mesh);
	if (rc) return rc;
	const bool f64 = opts.precision == GP_PREC_F64;
	DeviceMesh *dm = mesh->dev;
	VertexTables mt{};
	uint32_t S = 0;
	if ((rc = master_tables(*mesh, f64, mt, S))) return rc;
	if (S > MAXS) return fail(GP_ERR_CAPACITY, "plan: a fan needs %u slots", S);
	auto p = std::make_unique<gp_plan>();
	p->mesh = mesh;
	p->opts = opts;
	p->nv_orig = mesh->nv;
	p->ns = ns;
	p->dm = dm;
	if ((rc = plan_stream(p.get()))) return rc;
	// Rings of the sources on the device.
	Levels lv;
	if ((rc = lv.create(dm, p->stream))) return rc;
	std::vector<uint32_t> srcv(src, src + ns);
	uint32_t *d_srcv = nullptr;
	if ((rc = upload(d_srcv, srcv))) return rc;
	std::unique_ptr<uint32_t, hipError_t (*)(void *)> src_guard(d_srcv, hipFree);
	lv.clear();
	lv.seed(d_srcv, 0, ns);
	if ((rc = lv.expand())) return rc;
	if (dhat) {
		// Bands of dhat in place of the BFS levels of the reached vertices.
		if (!(width > 0)) return fail(GP_ERR_INVALID, "band width must be positive");
		std::vector<double> dh(dhat, dhat + mesh->nv);
		double *d_dhat = nullptr;
		uint32_t *d_bmax = nullptr;
		if ((rc = upload(d_dhat, dh))) return rc;
		std::unique_ptr<double, hipError_t (*)(void *)> dhat_guard(d_dhat, hipFree);
		if (hipMalloc(&d_bmax, 4) != hipSuccess) return fail(GP_ERR_NOMEM, "plan: device allocation failed");
		std::unique_ptr<uint32_t, hipError_t (*)(void *)> bmax_guard(d_bmax, hipFree);
		const uint32_t gb = std::max<uint32_t>(1, std::min<uint32_t>(2048, (mesh->nv + 255) / 256));
		k_fill_u32<<<1, 64, 0, p->stream>>>(d_bmax, 1, 0);
		k_band_levels<<<gb, 256, 0, p->stream>>>(lv.level, d_dhat, mesh->nv, width, d_bmax);
		GP_HIP(hipGetLastError());
		uint32_t bmax = 0;
		GP_HIP(hipMemcpyAsync(&bmax, d_bmax, 4, hipMemcpyDeviceToHost, p->stream));
		GP_HIP(hipStreamSynchronize(p->stream));
		if (bmax >= mesh->nv) return fail(GP_ERR_CAPACITY, "plan: band index %u past the vertex count", bmax);
		lv.sort();
		k_compact_limits<<<1, 1024, 0, p->stream>>>(lv.limits, lv.nl, lv.hist);
		GP_HIP(hipGetLastError());
	} else {
		lv.sort();
	}
	uint32_t nl = 0;
	GP_HIP(hipMemcpyAsync(&nl, lv.nl, 4, hipMemcpyDeviceToHost, p->stream));
	GP_HIP(hipStreamSynchronize(p->stream));
	if (nl < 2 || nl > mesh->nv + 1) return fail(GP_ERR_KERNEL, "plan: bad ring count %u", nl);
	uint32_t nr = 0;
	GP_HIP(hipMemcpy(&nr, lv.limits + (nl - 1), 4, hipMemcpyDeviceToHost));
	p->nl = nl;
	p->nr = nr;
	p->n = nr;
	p->max_iter = opts.max_sweeps ? opts.max_sweeps : (opts.uncapped ? 64 * nl + 64 : (nl << 1));
	p->block = SWEEP_BLOCK;
	p->nchunks = (nr + 63) / 64;
	p->max_slots = S;
	p->sb = 0;                               // per-chunk slot rows (each ring sorted longest fans first)
	p->n_flagged = 0;
	p->g64 = !f64 && opts.fp32_geometry == GP_GEO_FP64;
	uint32_t rows = 0;
	GP_HIP(hipMemcpy(&rows, lv.coff + p->nchunks, 4, hipMemcpyDeviceToHost));
	if (rows > p->nchunks * S) return fail(GP_ERR_KERNEL, "plan: bad chunk slot rows %u", rows);
	const size_t lanes = (size_t) rows * 64;
	if (hipMalloc(&p->d_limits, 4ull * nl) != hipSuccess || hipMalloc(&p->d_src, 4ull * ns) != hipSuccess ||
	    hipMalloc(&p->d_inv, 4ull * mesh->nv) != hipSuccess ||
	    hipMalloc(&p->d_coff, 4ull * (p->nchunks + 1)) != hipSuccess ||
	    hipMalloc(&p->d_cid, 4 * std::max<size_t>(lanes, 1)) != hipSuccess)
		return fail(GP_ERR_NOMEM, "plan: device allocation failed");
	if (f64) {
		if (hipMalloc(&p->d_cx64, sizeof(Edge64) * std::max<size_t>(lanes, 1)) != hipSuccess)
			return fail(GP_ERR_NOMEM, "plan: device allocation failed");
	} else if (hipMalloc(&p->d_cgeo, sizeof(float4) * std::max<size_t>(lanes, 1)) != hipSuccess ||
	           hipMalloc(&p->d_cgeo64, sizeof(double4) * std::max<size_t>(lanes, 1)) != hipSuccess) {
		return fail(GP_ERR_NOMEM, "plan: device allocation failed");
	}
	GP_HIP(hipMemcpyAsync(p->d_limits, lv.limits, 4ull * nl, hipMemcpyDeviceToDevice, p->stream));
	GP_HIP(hipMemcpyAsync(p->d_coff, lv.coff, 4ull * (p->nchunks + 1), hipMemcpyDeviceToDevice, p->stream));
	const uint32_t g = std::max<uint32_t>(1, std::min<uint32_t>(2048, (mesh->nv + 255) / 256));
	k_fill_u32<<<g, 256, 0, p->stream>>>(p->d_inv, mesh->nv, GP_NIL);
	k_positions<<<g, 256, 0, p->stream>>>(lv.sorted, nr, p->d_inv, p->d_src, ns);
	const uint32_t npos = p->nchunks * 64;
	const uint32_t gg = std::max<uint32_t>(1, std::min<uint32_t>(4096, (npos + 255) / 256));
	const uint32_t *nr_at = lv.limits + (nl - 1);
	if (f64)
		k_gather_tables<true, false><<<gg, 256, 0, p->stream>>>(lv.sorted, p->d_inv, nr_at, npos, S, lv.coff, mt,
		                                                         p->d_cid, nullptr, nullptr, nullptr, p->d_cx64);
	else
		k_gather_tables<false, false><<<gg, 256, 0, p->stream>>>(lv.sorted, p->d_inv, nr_at, npos, S, lv.coff, mt,
		                                                          p->d_cid, p->d_cgeo, p->d_cgeo64, nullptr, nullptr);
	GP_HIP(hipGetLastError());
	GP_HIP(hipStreamSynchronize(p->stream));
	if ((rc = sweep_grid(p.get(), f64))) return rc;
	if ((rc = plan_run_buffers(p.get(), f64))) return rc;
	*out = p.release();
	return GP_OK;
}

static KArgs make_args(const gp_plan *p)
{
	KArgs a{};
	a.limits = p->d_limits; a.coff = p->d_coff; a.sb = p->sb; a.cid = p->d_cid; a.cgeo = p->d_cgeo; a.cgeo64 = p->d_cgeo64;
	a.cx64 = p->d_cx64; a.fcum = p->d_fcum; a.order = p->d_order; a.dist0 = p->d_dist0; a.dist1 = p->d_dist1;
	a.st = p->d_state; a.trace = p->d_trace;
	a.tol = p->opts.tol; a.h = p->dm->h; a.inv_h = p->dm->inv_h;
	a.nl = p->nl; a.max_iter = p->max_iter; a.k2 = p->opts.uncapped ? 0u : 1u;
	a.cl0 = p->d_cl0; a.cl1 = p->d_cl1;
	return a;
}

// Enqueues sweeps k0 .. k0+count-1 (their k & 3 slots).
static void enqueue_sweeps(const gp_plan *p, const KArgs &a, uint32_t k0, uint32_t count)
{
	for (uint32_t q = 0; q < count; q++) {
		const uint32_t k4 = (k0 + q) & 3u;
		const bool cl = a.cl0 != nullptr;
		if (p->opts.precision == GP_PREC_F64) {
			if (cl && p->ms8) k_sweep<true, true, 8><<<p->grid, p->block, 0, p->stream>>>(a, k4);
			else if (cl) k_sweep<true, true, MAXS><<<p->grid, p->block, 0, p->stream>>>(a, k4);
			else if (p->ms8) k_sweep<true, false, 8><<<p->grid, p->block, 0, p->stream>>>(a, k4);
			else k_sweep<true, false, MAXS><<<p->grid, p->block, 0, p->stream>>>(a, k4);
		} else {
			if (p->g64) {
				if (cl && p->ms8) k_sweep<false, true, 8, true><<<p->grid, p->block, 0, p->stream>>>(a, k4);
				else if (cl) k_sweep<false, true, MAXS, true><<<p->grid, p->block, 0, p->stream>>>(a, k4);
				else if (p->ms8) k_sweep<false, false, 8, true><<<p->grid, p->block, 0, p->stream>>>(a, k4);
				else k_sweep<false, false, MAXS, true><<<p->grid, p->block, 0, p->stream>>>(a, k4);
			} else if (cl && p->ms8) k_sweep<false, true, 8><<<p->grid, p->block, 0, p->stream>>>(a, k4);
			else if (cl) k_sweep<false, true, MAXS><<<p->grid, p->block, 0, p->stream>>>(a, k4);
			else if (p->ms8) k_sweep<false, false, 8><<<p->grid, p->block, 0, p->stream>>>(a, k4);
			else k_sweep<false, false, MAXS><<<p->grid, p->block, 0, p->stream>>>(a, k4);
		}
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
	if (p->pp) return patch_run(p->pp, stats);
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
	if (p->d_cl0) {
		k_cl_init<<<ig, 256, 0, p->stream>>>(p->d_cl0, p->d_cl1, p->n);
		k_cl_sources<<<(p->ns + 255) / 256, 256, 0, p->stream>>>(p->d_cl0, p->d_cl1, p->d_src, p->ns);
	}
	GP_HIP(hipGetLastError());
	GP_HIP(hipEventRecord(p->e0, p->stream));
	// Replays until the device reports done, two replays in flight: after
	// each replay the state is copied to its pinned snapshot and an event is
	// recorded; the host waits for the older replay's event while the newer
	// one runs, so the GPU does not idle on the poll.  The device time ends
	// at the event of the replay whose state says done (a replay queued
	// behind it only meets done slots and returns at once).  One sweep more
	// than the sweep count runs (it records the last sweep's outcome).  j
	// grows by one ring per sweep until |limits|-1, so at least nl-1-j sweeps
	// remain: big graphs while that bound allows, small ones near the end.
	uint32_t launched = 0, j_seen = 2;
	const uint32_t cap = p->max_iter;
	auto replay = [&](uint32_t s) -> int {
		const bool big = p->nl - 1 > j_seen + BIG + (launched ? BIG : 0);
		const uint32_t cnt = big ? BIG : SMALL;
		if (graph) GP_HIP(hipGraphLaunch(big ? p->g_big : p->g_small, p->stream));
		else enqueue_sweeps(p, a, launched, cnt);
		launched += cnt;
		GP_HIP(hipGetLastError());
		GP_HIP(hipMemcpyAsync(p->h_state + s, p->d_state, sizeof(PState), hipMemcpyDeviceToHost, p->stream));
		GP_HIP(hipEventRecord(p->ev[s], p->stream));
		return GP_OK;
	};
	uint32_t cur = 0;
	bool done = done0 != 0;
	if (!done) {
		int rc = replay(0);
		if (rc) return rc;
		for (;;) {
			// Queue the next replay (if the cap allows), then wait for the
			// older one; the replay queued here is the one checked next.
			const bool queued = launched <= cap;
			if (queued && (rc = replay(cur ^ 1u))) return rc;
			GP_HIP(hipEventSynchronize(p->ev[cur]));
			if (p->h_state[cur].done) { done = true; break; }
			if (!queued) break;
			// Graphs hold a multiple of 4 sweeps: slot 3 is a replay's last sweep.
			j_seen = p->h_state[cur].slot[3].j;
			cur ^= 1u;
		}
	}
	GP_HIP(hipStreamSynchronize(p->stream));
	if (!done) return fail(GP_ERR_KERNEL, "ptp: schedule did not finish in %u sweeps", launched);
	if (done0) {
		GP_HIP(hipEventRecord(p->e1, p->stream));
		GP_HIP(hipMemcpyAsync(p->h_state + cur, p->d_state, sizeof(PState), hipMemcpyDeviceToHost, p->stream));
		GP_HIP(hipStreamSynchronize(p->stream));
	}
	if (cur != 0) std::memcpy(p->h_state, p->h_state + cur, sizeof(PState));
	float ms = 0;
	GP_HIP(hipEventElapsedTime(&ms, p->e0, done0 ? p->e1 : p->ev[cur]));
	p->has_run = true;
	if (stats) {
		stats->sweeps = p->h_state->sweeps;
		stats->n_limits = p->nl;
		stats->relaxations = p->h_state->relax;
		stats->cond_checks = p->h_state->cond;
		stats->f64_updates = p->h_state->f64;
		stats->patch_visits = 0;
		stats->deferred = 0;
		stats->kernel_ms = ms;
		stats->total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
	}
	return GP_OK;
}

int plan_result(gp_plan *p, double *dist)
{
	if (!p || !dist) return fail(GP_ERR_INVALID, "plan_result: null argument");
	if (p->pp) return patch_result(p->pp, dist);
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
	if (p->d_inv) {
		double *d_orig = nullptr;
		if (hipMalloc(&d_orig, sizeof(double) * p->nv_orig) != hipSuccess)
			return fail(GP_ERR_NOMEM, "plan_result: allocation failed");
		const uint32_t go = std::max<uint32_t>(1, std::min<uint32_t>(1024, (p->nv_orig + 255) / 256));
		k_from_positions<<<go, 256, 0, p->stream>>>(p->d_out, p->d_inv, p->nv_orig, d_orig);
		const hipError_t e1 = hipMemcpyAsync(dist, d_orig, sizeof(double) * p->nv_orig, hipMemcpyDeviceToHost, p->stream);
		const hipError_t e2 = hipStreamSynchronize(p->stream);
		(void) hipFree(d_orig);
		if (e1 != hipSuccess || e2 != hipSuccess) return fail(GP_ERR_HIP, "plan_result: copy failed");
	} else if (p->coal) {
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

int plan_clusters(gp_plan *p, uint32_t *out)
{
	if (!p || !out) return fail(GP_ERR_INVALID, "plan_clusters: null argument");
	if (p->pp || !p->d_cl0) return fail(GP_ERR_INVALID, "plan_clusters: the plan was made without clusters");
	if (!p->has_run) return fail(GP_ERR_INVALID, "plan_clusters: plan has not run");
	const uint32_t d = p->h_state->final_d;
	const bool last_read = p->opts.result_buffer == GP_RESULT_LAST_READ;
	const uint32_t *buf = (last_read ? !d : d) ? p->d_cl1 : p->d_cl0;
	std::vector<uint32_t> tmp(p->n);
	GP_HIP(hipMemcpyAsync(tmp.data(), buf, 4ull * p->n, hipMemcpyDeviceToHost, p->stream));
	GP_HIP(hipStreamSynchronize(p->stream));
	if (p->coal) {
		for (uint32_t v = 0; v < p->nv_orig; v++) out[v] = p->inv[v] != NIL ? tmp[p->inv[v]] : GP_NIL;
	} else {
		std::memcpy(out, tmp.data(), 4ull * p->n);
	}
	return GP_OK;
}

int plan_trace(gp_plan *p, uint32_t *out, uint32_t cap, uint32_t *n)
{
	if (p && p->pp) return patch_trace(p->pp, out, cap, n);
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

int plan_create_patches(Che *mesh, const uint32_t *src, uint32_t ns, const double *dhat, const gp_ptp_opts &opts,
                        gp_plan **out)
{
	*out = nullptr;
	PatchPlan *pp = nullptr;
	int rc = patch_create(mesh, src, ns, dhat, opts, &pp);
	if (rc) return rc;
	auto p = std::make_unique<gp_plan>();
	p->pp = pp;
	p->mesh = mesh;
	p->opts = opts;
	*out = p.release();
	return GP_OK;
}

// compute_error (src/test_geodesics_ptp.cpp:362-371) on the device: the sum of
// |d - exact| / exact over exact > 0 (the caller scales by 100 / (n - s)).
template <typename T>
__global__ void k_iter_err(const T *d, const double *exact, uint32_t n, double *acc)
{
	double e = 0;
	for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x)
		if (exact[v] > 0) e += fabs((double) d[v] - exact[v]) / exact[v];
	for (uint32_t o = 32; o > 0; o >>= 1) e += __shfl_xor(e, o);
	if ((threadIdx.x & 63u) == 0) atomicAdd(acc, e);
}

int ptp_iter_error(Che *mesh, const uint32_t *src, uint32_t ns, const double *exact, bool f64, uint32_t *iter,
                   double *err, uint32_t cap, uint32_t *n_err, uint32_t *band, uint32_t cap_band, uint32_t *n_band)
{
	if (!mesh || !src || !exact || !n_err || !n_band) return fail(GP_ERR_INVALID, "iter_error: null argument");
	*n_err = *n_band = 0;
	const uint32_t nv = mesh->nv;
	std::vector<uint32_t> lev(nv), sorted(nv), lim;
	int rc = toplesets_host(*mesh, src, ns, lev.data(), sorted.data(), lim);
	if (rc) return rc;
	if (lim.size() < 2 || lim.back() != nv)
		return fail(GP_ERR_INVALID, "iter_error: every vertex must be reached from the sources");
	gp_ptp_opts o;
	gp_ptp_opts_default(&o);
	o.precision = f64 ? GP_PREC_F64 : GP_PREC_F32;
	o.coalesce = 1;
	o.uncapped = 1;
	o.launch_mode = GP_LAUNCH_STREAM;
	gp_plan *pp = nullptr;
	if ((rc = plan_create_rings(mesh, src, ns, lim.data(), (uint32_t) lim.size(), sorted.data(), o, &pp))) return rc;
	std::unique_ptr<gp_plan> p(pp);
	// Exact distances in the plan's ring numbering (test_geodesics_ptp_coalescence.cu:37).
	std::vector<double> exc(nv);
	for (uint32_t v = 0; v < nv; v++) exc[p->inv[v]] = exact[v];
	double *d_exact = nullptr, *d_acc = nullptr;
	struct Free {
		std::vector<void *> v;
		~Free() { for (void *q : v) (void) hipFree(q); }
	} fr;
	if (hipMalloc(&d_exact, 8ull * nv) != hipSuccess || hipMalloc(&d_acc, 8) != hipSuccess) {
		fr.v = { d_exact, d_acc };
		return fail(GP_ERR_NOMEM, "iter_error: allocation failed");
	}
	fr.v = { d_exact, d_acc };
	GP_HIP(hipMemcpy(d_exact, exc.data(), 8ull * nv, hipMemcpyHostToDevice));
	const uint32_t ig = std::max<uint32_t>(1, std::min<uint32_t>(1024, (p->n + 255) / 256));
	const uint32_t done0 = p->nl < 3 ? 1u : 0u;
	if (f64) {
		k_init<double><<<ig, 256, 0, p->stream>>>((double *) p->d_dist0, (double *) p->d_dist1, p->n, p->d_state, done0);
		k_sources<double><<<(p->ns + 255) / 256, 256, 0, p->stream>>>((double *) p->d_dist0, (double *) p->d_dist1, p->d_src, p->ns);
	} else {
		k_init<float><<<ig, 256, 0, p->stream>>>((float *) p->d_dist0, (float *) p->d_dist1, p->n, p->d_state, done0);
		k_sources<float><<<(p->ns + 255) / 256, 256, 0, p->stream>>>((float *) p->d_dist0, (float *) p->d_dist1, p->d_src, p->ns);
	}
	GP_HIP(hipGetLastError());
	const KArgs a = make_args(p.get());
	uint32_t ne = 0, nb = 0;
	for (uint32_t k = 0; !done0 && k <= p->max_iter; k++) {
		enqueue_sweeps(p.get(), a, k, 1);
		GP_HIP(hipGetLastError());
		GP_HIP(hipMemcpyAsync(p->h_state, p->d_state, sizeof(PState), hipMemcpyDeviceToHost, p->stream));
		GP_HIP(hipStreamSynchronize(p->stream));
		if (p->h_state->done) break;
		// Sweep k + 1 just ran with this state (the uncapped loop has no K/2).
		const Slot s = p->h_state->slot[k & 3u];
		if (nb < cap_band && band) {
			uint32_t *r = band + 4 * (size_t) nb;
			r[0] = k + 1; r[1] = s.i; r[2] = s.j; r[3] = lim[s.j] - lim[s.i];
		}
		nb++;
		if (s.j == p->nl - 1) {
			const void *nd = s.d ? p->d_dist0 : p->d_dist1;    // the buffer the sweep wrote
			GP_HIP(hipMemsetAsync(d_acc, 0, 8, p->stream));
			if (f64) k_iter_err<double><<<ig, 256, 0, p->stream>>>((const double *) nd, d_exact, nv, d_acc);
			else k_iter_err<float><<<ig, 256, 0, p->stream>>>((const float *) nd, d_exact, nv, d_acc);
			GP_HIP(hipGetLastError());
			double acc = 0;
			GP_HIP(hipMemcpyAsync(&acc, d_acc, 8, hipMemcpyDeviceToHost, p->stream));
			GP_HIP(hipStreamSynchronize(p->stream));
			if (ne < cap && iter && err) { iter[ne] = k + 1; err[ne] = acc * 100.0 / (double) (nv - ns); }
			ne++;
		}
	}
	*n_err = ne;
	*n_band = nb;
	return GP_OK;
}

}  // namespace gp

extern "C" void gp_plan_destroy(gp_plan *p) { delete p; }
