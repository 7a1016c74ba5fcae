// The window sweep kernel of the parallel toplesets propagation and its
// device-resident schedule state, shared by the solver plans (ptp_gpu.hip)
// and any other driver of the same schedule.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "device.hpp"
#include "sweep_dev.hpp"
#include "update_step.hpp"

namespace gp {

// ---------------------------------------------------------------------------
// Kernels

// Schedule state in device memory.  Sweep kernel k of a run (k counted from
// 0) reads the state of sweep k-1 from slot[(k-1) & 3] together with its
// non-converged count bad[(k-1) & 3], advances it (i, j, d as the reference
// loop does), publishes it in slot[k & 3] (workgroup 0), zeroes bad[(k+1) & 3]
// for the next sweep, and relaxes its own window counting into bad[k & 3].
// No workgroup reads a slot or counter while another workgroup of the same
// launch writes it, so kernel boundaries are the only synchronisation.  The
// graphs hold a multiple of 4 sweeps so k & 3 stays continuous.
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
	uint32_t done, sweeps, final_d, check_bad;   // check_bad: count of a check sweep (band plans)
	unsigned long long relax, cond, f64;
};

struct KArgs {
	const uint32_t *limits;
	const uint32_t *coff;      // per chunk: first slot row
	uint32_t sb;               // > 0: every chunk has sb slots (chunk c at row c * sb; coff unused)
	const uint32_t *cid;
	const float2 *cgeo;        // FP32 geometry: q0, g per slot
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
	uint32_t npos;             // positions of the plan (check sweeps cover [0, npos))
	uint32_t xblk;             // log2 of the chunk blocks dealt to XCDs (0: contiguous eighths)
	uint32_t k2;               // 1: K/2 limit (the drivers); 0: the uncapped iter_error loop
	uint32_t *cl0, *cl1;       // cluster labels (double-buffered like the distances), or null
	const uint8_t *closed;     // vertex-major sweeps: closed-fan flags of the mesh
};

// Relaxes the window [start, end) (positions); returns the wave's count of
// non-converged vertices among positions [start, cend).  VM: the tables are
// the mesh's vertex-major ones (slot k of vertex v at coff[v] + k, coff = the
// fan offsets) and position p relaxes vertex order[p]; positions and vertices
// at or past npos (the vertex count) are skipped.  CHK (the check sweeps): a
// value that did not change counts as converged (a source's 0 / 0 and an
// unreached INF, which the ring test of the reference's loop never meets:
// its window starts past the sources).
template <bool F64, bool CL, uint32_t MS, bool G64, bool VM = false, bool CHK = false>
__device__ __forceinline__ uint32_t sweep_body(const KArgs &a, uint32_t start, uint32_t end, uint32_t cend,
                                               const typename std::conditional<F64, double, float>::type *od,
                                               typename std::conditional<F64, double, float>::type *nd)
{
	using T = typename std::conditional<F64, double, float>::type;
	if (start >= end) return 0;
	const uint32_t c0 = start >> 6, c1 = (end + 63u) >> 6;
	const uint32_t lane = threadIdx.x & 63u;
	// XCD-aware split: workgroups b with b % 8 == x run on XCD x.  xblk == 0:
	// XCD x takes the x-th contiguous eighth of the window's chunks.  xblk > 0:
	// chunks are dealt to XCDs in blocks of 2^xblk, block q to XCD q % 8, so a
	// chunk is relaxed on the same XCD in every sweep (its tables and its
	// neighbours' values stay in that XCD's L2).
	const uint32_t x = blockIdx.x & 7u;
	const uint32_t nbx = (gridDim.x - x + 7u) >> 3;
	const uint32_t wpb = blockDim.x >> 6;
	// The wave's index, taken from the first lane: the same for the whole
	// wave, and known to the compiler as such, so the chunk index, its slot
	// rows and the slot count live in scalar registers and the per-slot tests
	// are scalar branches instead of per-lane compares and exec masks.
	const uint32_t lw = (blockIdx.x >> 3) * wpb + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
	const uint32_t nw_x = nbx * wpb;
	uint32_t bad = 0;
	static_assert(!(VM && CL), "vertex-major sweeps carry no labels");
	auto chunk = [&](uint32_t c) {
		const uint32_t pos = (c << 6) + lane;
		bool lbad = false;
		if constexpr (VM) {
			const uint32_t v = pos < a.npos ? a.order[pos] : GP_NIL;
			if (pos >= start && pos < end && v < a.npos) {
				const T old = od[v];
				const uint32_t b0 = a.coff[v], nf = a.coff[v + 1] - b0;
				const uint32_t S = (nf && a.closed[v]) ? nf - 1 : nf;
				const T best = S ? relax_lane<F64, 1, true, false, MS, G64>(a, (size_t) b0, S, od, old) : old;
				nd[v] = best;
				if (pos < cend) lbad = !(fabs((double) best - (double) old) / (double) old < a.tol);
				if (CHK && best == old) lbad = false;
			}
			bad += (uint32_t) __popcll(__ballot(lbad));
			return;
		}
		// Both chunk offsets are requested together, in one block: written as
		// two selects on sb, the second load was issued only after the first
		// had returned, one more round trip in the chain every sweep waits on.
		uint32_t cb = c * a.sb, S = a.sb;
		if (!a.sb) {
			const uint32_t e0 = a.coff[c], e1 = a.coff[c + 1];
			cb = e0;
			S = e1 - e0;
		}
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
				uint32_t cl = ocl[v];
				if (bk != NO_TRI) {
					const uint32_t ia = a.cid[base + 64 * (size_t) bk] & SLOT_ID_MASK;
					const uint32_t nb = bk + 1 < S ? a.cid[base + 64 * (size_t) (bk + 1)] : SLOT_EMPTY;
					const uint32_t ib = ((nb & SLOT_EMPTY) ? a.cid[base] : nb) & SLOT_ID_MASK;
					cl = od[ib] < od[ia] ? ocl[ib] : ocl[ia];
				}
				ncl[v] = cl;
			}
			if (pos < cend) lbad = !(fabs((double) best - (double) old) / (double) old < a.tol);
			if (CHK && best == old) lbad = false;
		}
		bad += (uint32_t) __popcll(__ballot(lbad));
	};
	if (a.xblk && gridDim.x >= 8u) {
		const uint32_t L = a.xblk, B = 1u << L;
		const uint32_t q0 = c0 >> L, q1 = (c1 - 1) >> L;
		const uint32_t qx = q0 + ((x + 8u - (q0 & 7u)) & 7u);     // XCD x's first block
		for (uint32_t k = lw;; k += nw_x) {
			const uint32_t q = qx + 8u * (k >> L);
			if (q > q1) break;
			const uint32_t c = (q << L) + (k & (B - 1u));
			if (c >= c0 && c < c1) chunk(c);
		}
	} else {
		const uint32_t ngrp = gridDim.x < 8u ? gridDim.x : 8u;
		const uint32_t len = c1 - c0;
		const uint32_t part = (len + ngrp - 1) / ngrp;
		const uint32_t lo = c0 + x * part;
		const uint32_t hi = min(c1, lo + part);
		for (uint32_t c = lo + lw; c < hi; c += nw_x) chunk(c);
	}
	return bad;
}

// One wave per workgroup: waves of a sweep finish and free their slots
// independently, and the XCD split deals the window out in finer pieces
// (measured on the 1M field against 4 waves per workgroup: 2.5 % faster on
// the reference schedule, 4 % on the learned bands).
constexpr uint32_t SWEEP_BLOCK = 64;

// Sweep k (k4 = k & 3) of a run: advance the schedule, then relax its window.
// k4 = 4 + d: a check sweep of the band schedules instead -- one Jacobi sweep
// over every position [0, npos) from buffer d into the other, counting into
// check_bad the positions whose relative change is not below tol (an INF
// value next to a finite neighbour counts).
// The state pointer and k4 come first, as plain arguments, so that they can
// be preloaded into scalar registers at wave launch: the schedule slot's load
// then does not wait for the kernel-argument segment (a.st is the same
// pointer).
template <bool F64, bool CL, uint32_t MS, bool G64 = false, bool VM = false>
__global__ __launch_bounds__(256) void k_sweep(PState *st, uint32_t k4, KArgs a)
{
	using T = typename std::conditional<F64, double, float>::type;
	if (k4 >= 4u) {
		const uint32_t d = k4 & 1u;
		const T *cod = (const T *) (d ? a.dist1 : a.dist0);
		T *cnd = (T *) (d ? a.dist0 : a.dist1);
		const uint32_t cbad = sweep_body<F64, CL, MS, G64, VM, true>(a, 0, a.npos, a.npos, cod, cnd);
		if ((threadIdx.x & 63u) == 0 && cbad) atomicAdd(&st->check_bad, cbad);
		return;
	}
	const uint32_t kp = (k4 + 3u) & 3u;
	const Slot prev = st->slot[kp];
	// Read with the slot, before any store of this launch, and held in a
	// scalar register here (the empty asm takes it as an input, so the load
	// is not sunk into the branch below): both loads are issued together and
	// the count does not add a second round trip to the chain every sweep
	// starts with.
	const uint32_t nbad = st->bad[kp];
	asm volatile("" ::"s"(nbad));
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
	const uint32_t bad = sweep_body<F64, CL, MS, G64, VM>(a, start, end, cend, od, nd);
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

}  // namespace gp
