// Dataflow form of the window sweep (FP32 ring and band plans): one
// persistent cooperative launch replaces the launch per sweep.  Wave w owns
// the chunks c = w + q * waves (q < 32) and runs the window schedule sweep by
// sweep over them; a chunk starts sweep k as soon as the neighbour values it
// reads (those of sweep k-1) are published, instead of waiting at a kernel
// boundary for the whole previous sweep.
//
// Schedule and arithmetic are those of k_sweep (the reference loop,
// src/geodesics_ptp.cpp:61-109, src/cuda/geodesics_ptp.cu:198-231): the same
// window [i', j) per sweep, the same ring test, the same double buffer and the
// same relaxation (relax_lane); only the synchronisation differs, so fields,
// sweep counts and statistics are bit-identical.
//
// Publication (agent-scope atomic loads and stores: coherent across the XCD
// L2s without cache maintenance, so the read-only slot tables stay cached):
//  * distances are (value, tag) pairs, tag = 1 + the sweep that wrote them;
//  * cflag[c] = k + 1 once chunk c's writes of sweep k have completed, FROZEN
//    once the window start has passed the chunk;
//  * hist[k] = the window start i'_k of sweep k (| DONE), published in
//    order (pub = states published) by the last chunk of sweep k-1's tested
//    ring to finish, which also adds that sweep's statistics.
//
// What chunk c needs at sweep k:
//  * its window: j_k = min(2 + k, |limits| - 1) is fixed; i' grows by at most
//    one ring per sweep, so from the latest published state m,
//    i'_k <= max(i'_m, j_k / 2) + (k - m): a chunk starting at or past that
//    ring + 1 is inside the window and not ring-tested, and waits for no
//    state; a chunk below it waits for state k;
//  * each neighbour value of sweep k-1 (buffer k & 1): a neighbour at or past
//    the window end of k-1 holds its initial value; one inside the window of
//    k-1 is ready when its tag is k (its next write to that buffer, at sweep
//    k+1, reads this chunk's value of sweep k, so it cannot come early); one
//    below the window start of k-1 is frozen and ready once its chunk has
//    completed sweep k-2.
// A wave processes (sweep, chunk) items in sweep order and every dependency
// of an item of sweep k is an item of sweep k-1 or state k, so with every
// wave resident (cooperative launch) the grid cannot deadlock.  Every wait is
// still bounded in wall-clock time: past the deadline the waiting wave sets
// abort, every wave leaves at its next wait, and the host reports the run as
// failed.
#pragma once

#include <hip/hip_runtime.h>

#include "sweep_kernel.hpp"

namespace gp {

constexpr uint32_t FLOW_DONE = 0x80000000u;     // hist: the loop ended with this state
constexpr uint32_t FLOW_IMASK = 0x7FFFFFFFu;
constexpr uint32_t FLOW_FROZEN = 0xFFFFFFFFu;   // cflag of a chunk below the window start
constexpr uint32_t FLOW_NEVER = 0xFFFFFFFFu;
constexpr uint32_t FLOW_MAX_OWN = 32;           // chunks per wave

struct FlowState {
	uint32_t pub;     // states published: hist[0 .. pub-1]
	uint32_t abort;   // a wait passed the deadline
	uint32_t front;   // largest sweep in which a chunk entering the window has finished
	uint32_t pad;
};

struct FlowArgs {
	KArgs a;
	unsigned long long *pr0, *pr1;   // (value, tag) pairs of the two distance buffers
	uint32_t *cflag;                 // per chunk
	uint32_t *hist;                  // per sweep
	unsigned long long *acc;         // per sweep: (tested chunks finished << 32) | non-converged count
	FlowState *fs;
	uint32_t nchunks;
	uint32_t nr;                     // positions in rings: limits[|limits| - 1]
	unsigned long long budget;       // wall-clock ticks (100 MHz) a run may take
};

__device__ __forceinline__ uint32_t ld_agent(const uint32_t *p)
{
	return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned long long ld_agent(const unsigned long long *p)
{
	return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// The same load of a word every lane reads, its value kept in a scalar register.
__device__ __forceinline__ uint32_t ld_agent_uniform(const uint32_t *p)
{
	return (uint32_t) __builtin_amdgcn_readfirstlane((int) ld_agent(p));
}
__device__ __forceinline__ void st_agent(uint32_t *p, uint32_t v)
{
	__hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_agent(unsigned long long *p, unsigned long long v)
{
	__hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Waits until this wave's earlier memory operations (agent-scope stores
// included) have completed; a store issued after it is seen after them.
__device__ __forceinline__ void flow_wait_stores()
{
	__asm__ volatile("s_waitcnt vmcnt(0)" ::: "memory");
}

// j of sweep k (j grows by one ring per sweep from 2 up to |limits| - 1;
// the runs have |limits| >= 3).
__device__ __forceinline__ uint32_t flow_j(uint32_t k, uint32_t nl)
{
	return 2u + k < nl - 1u ? 2u + k : nl - 1u;
}

// True when the run must stop: abort set by another wave (checked every 8th
// call), or this wave passed the deadline (then it sets abort).
__device__ __forceinline__ bool flow_expired(const FlowArgs &f, unsigned long long deadline, uint32_t it)
{
	if ((it & 7u) == 0 && ld_agent_uniform(&f.fs->abort)) return true;
	if (wall_clock64() > deadline) {
		st_agent(&f.fs->abort, 1u);
		return true;
	}
	return false;
}

// Neighbour distances of sweep kk: the pairs of buffer kk & 1 and what is
// needed to tell when each one is final.
struct FlowOD {
	const unsigned long long *pr;
	const uint32_t *cflag;
	const FlowArgs *f;
	uint32_t kk, c;
	uint32_t ep;          // window end of sweep kk-1 (0 for kk = 0: every value is initial)
	uint32_t sp_lo, sp_hi;// window start of sweep kk-1 is in [sp_lo, sp_hi]
	unsigned long long deadline;
	mutable bool failed = false;
	__device__ float operator[](uint32_t i) const { return __uint_as_float((uint32_t) ld_agent(pr + i)); }
};

// The exact window start of sweep k-1 (state k-1 published): waits for it.
__device__ __forceinline__ bool flow_state(const FlowArgs &f, uint32_t k, unsigned long long deadline, uint32_t &h)
{
	for (uint32_t it = 0;; it++) {
		const uint32_t P = ld_agent_uniform(&f.fs->pub);
		if (P > k) {
			h = ld_agent_uniform(f.hist + k);
			return true;
		}
		const uint32_t hl = ld_agent_uniform(f.hist + (P - 1));
		if (hl & FLOW_DONE) {   // the loop ended before sweep k
			h = hl;
			return true;
		}
		if (flow_expired(f, deadline, it)) return false;
		__builtin_amdgcn_s_sleep(1);
	}
}

// Class of neighbour position p for the sweep's readiness test: 0 ready
// (initial value, or this wave's own write), 1 written in sweep kk-1 (ready
// when its tag is kk), 2 frozen (ready once its chunk finished sweep kk-2),
// 3 the window start of kk-1 is needed to tell.
__device__ __forceinline__ uint32_t flow_class(const FlowOD &od, uint32_t p, uint32_t sp_lo, uint32_t sp_hi)
{
	if (p >= od.ep || (p >> 6) == od.c) return 0;
	if (p >= sp_hi) return 1;
	if (p < sp_lo) return 2;
	return 3;
}

// Loads the pairs (and, for frozen neighbours, their chunks' flags) of the
// slots not yet ready, all issued before any is tested; returns the updated
// ready mask, and in amb the slots whose class needs the exact window start.
template <uint32_t MS>
__device__ __forceinline__ uint32_t flow_poll(const FlowOD &od, const uint32_t (&id)[MS], uint32_t S, float (&t)[MS],
                                              uint32_t ok, uint32_t sp_lo, uint32_t sp_hi, uint32_t &amb)
{
	unsigned long long x[MS];
	uint32_t fl[MS], cls[MS];
#pragma unroll
	for (uint32_t k = 0; k < MS; k++) {
		const uint32_t p = id[k] & SLOT_ID_MASK;
		cls[k] = flow_class(od, p, sp_lo, sp_hi);
		if (k < S && !((ok >> k) & 1u)) {
			x[k] = ld_agent(od.pr + p);
			// Every lane loads a flag; lanes without a frozen neighbour
			// read chunk 0's (one request for the wave).
			fl[k] = ld_agent(od.cflag + (cls[k] == 2 ? (p >> 6) : 0u));
		}
	}
	amb = 0;
#pragma unroll
	for (uint32_t k = 0; k < MS; k++) {
		if (k < S && !((ok >> k) & 1u)) {
			t[k] = __uint_as_float((uint32_t) x[k]);
			const bool r = cls[k] == 0 || (cls[k] == 1 && (uint32_t) (x[k] >> 32) >= od.kk) ||
			               (cls[k] == 2 && (fl[k] == FLOW_FROZEN || fl[k] + 1u >= od.kk));
			ok |= r ? (1u << k) : 0u;
			amb |= cls[k] == 3 ? (1u << k) : 0u;
		}
	}
	return ok;
}

template <uint32_t MS>
__device__ __forceinline__ void gather_slots(const FlowOD &od, const uint32_t (&id)[MS], uint32_t S, float (&t)[MS])
{
	uint32_t ok = 0, amb = 0;
	uint32_t sp_lo = od.sp_lo, sp_hi = od.sp_hi;
#pragma unroll
	for (uint32_t k = 0; k < MS; k++) {
		t[k] = INFINITY;
		if (k >= S) ok |= 1u << k;
	}
	ok = flow_poll(od, id, S, t, ok, sp_lo, sp_hi, amb);
	for (uint32_t it = 0; !__all(ok == (1u << MS) - 1u); it++) {
		if (__any(amb != 0)) {
			// Band plans in the bound mode: the exact window start of kk-1.
			uint32_t h = 0;
			if (!flow_state(*od.f, od.kk - 1u, od.deadline, h) || (h & FLOW_DONE)) {
				od.failed = true;
				return;
			}
			sp_lo = sp_hi = od.f->a.limits[h & FLOW_IMASK];
		} else {
			if (flow_expired(*od.f, od.deadline, it)) {
				od.failed = true;
				return;
			}
			__builtin_amdgcn_s_sleep(1);
		}
		ok = flow_poll(od, id, S, t, ok, sp_lo, sp_hi, amb);
	}
}

// The last tested chunk of sweep k: the reference loop tail (as k_sweep
// applies it at the start of the next sweep), the sweep's statistics and
// trace row, then state k+1.
__device__ __forceinline__ void flow_publish(const FlowArgs &f, uint32_t k, uint32_t ik, uint32_t ws, uint32_t wc,
                                             uint32_t we, uint32_t jk, uint32_t nbad)
{
	const KArgs &a = f.a;
	const uint32_t in = ik + (nbad == 0 ? 1u : 0u);
	const uint32_t jn = flow_j(k + 1u, a.nl);
	const uint32_t iter = k + 1u;
	const bool done = !(in < jn) || !(iter < a.max_iter);
	const uint32_t ipn = a.k2 && in < (jn >> 1) ? (jn >> 1) : in;
	PState *st = a.st;
	atomicAdd(&st->relax, (unsigned long long) (we - ws));
	atomicAdd(&st->cond, (unsigned long long) (wc - ws));
	if (a.fcum) atomicAdd(&st->f64, a.fcum[we] - a.fcum[ws]);
	if (a.trace) {
		uint32_t *r = a.trace + 8 * (size_t) k;
		r[0] = iter; r[1] = ik; r[2] = jk; r[3] = ws; r[4] = we; r[5] = wc - ws; r[6] = wc - ws - nbad; r[7] = 0;
	}
	if (done) {
		Slot &s = st->slot[3];
		s.i = in; s.j = jn; s.iter = iter; s.done = 1;
		st->sweeps = iter;
		st->final_d = iter & 1u;
		st->done = 1;
	}
	st_agent(f.hist + iter, ipn | (done ? FLOW_DONE : 0u));
	flow_wait_stores();
	st_agent(&f.fs->pub, iter + 1u);
}

enum : uint32_t { FLOW_SWEPT = 0, FLOW_OFF = 1, FLOW_STOP = 2 };

// Sweep k of chunk c (kin: the sweep in which it enters the window; pk: the
// entering sweep of the chunk that holds the previous ring's last position).
template <uint32_t MS, bool G64>
__device__ uint32_t flow_chunk(const FlowArgs &f, uint32_t c, uint32_t k, uint32_t kin, uint32_t pk,
                               unsigned long long deadline)
{
	const KArgs &a = f.a;
	const uint32_t lane = threadIdx.x & 63u;
	const uint32_t first = c << 6;
	const uint32_t last = min(first + 63u, f.nr - 1u);
	const uint32_t jk = flow_j(k, a.nl);
	const uint32_t we = a.limits[jk];
	// An entering chunk sleeps until the front is near (its wave may have
	// jumped far ahead of it): polling one word, the longer the farther.
	if (k == kin && pk > 0) {
		for (uint32_t it = 0;; it++) {
			const uint32_t fr = ld_agent_uniform(&f.fs->front);
			if (fr >= pk) break;
			if ((it & 15u) == 15u) {
				const uint32_t P = ld_agent_uniform(&f.fs->pub);
				if (ld_agent_uniform(f.hist + (P - 1)) & FLOW_DONE) return FLOW_STOP;
			}
			if (flow_expired(f, deadline, it)) return FLOW_STOP;
			for (uint32_t s = min(pk - fr, 24u); s > 0; s--) __builtin_amdgcn_s_sleep(32);
		}
	}
	// The window start of sweep k: exact, or bounded away from this chunk.
	bool exact = true;
	uint32_t ik = 0, im = 0, m = 0;
	{
		const uint32_t P = ld_agent_uniform(&f.fs->pub);
		uint32_t h;
		if (P > k) {
			h = ld_agent_uniform(f.hist + k);
		} else {
			m = P - 1u;
			h = ld_agent_uniform(f.hist + m);
			if (!(h & FLOW_DONE)) {
				im = h & FLOW_IMASK;
				if (a.limits[im] > last) {
					if (lane == 0) st_agent(f.cflag + c, FLOW_FROZEN);
					return FLOW_OFF;
				}
				const uint32_t u = max(im, a.k2 ? (jk >> 1) : 0u) + (k - m);
				exact = !(u + 1u < a.nl && a.limits[u + 1u] <= first);
				if (exact && !flow_state(f, k, deadline, h)) return FLOW_STOP;
			}
		}
		if (h & FLOW_DONE) return FLOW_STOP;
		if (exact) ik = h & FLOW_IMASK;
	}
	uint32_t ws = first, wc = 0;
	if (exact) {
		ws = a.limits[ik];
		if (ws > last) {
			if (lane == 0) st_agent(f.cflag + c, FLOW_FROZEN);
			return FLOW_OFF;
		}
		wc = a.limits[ik + 1u];
	}
	// Neighbour classes of sweep k-1.
	FlowOD od;
	od.pr = (k & 1u) ? f.pr1 : f.pr0;
	od.cflag = f.cflag;
	od.f = &f;
	od.kk = k;
	od.c = c;
	od.deadline = deadline;
	od.ep = 0;
	od.sp_lo = od.sp_hi = 0;
	if (k > 0) {
		const uint32_t jp = flow_j(k - 1u, a.nl);
		od.ep = a.limits[jp];
		if (exact) {
			od.sp_lo = od.sp_hi = a.limits[ld_agent_uniform(f.hist + (k - 1u)) & FLOW_IMASK];
		} else {
			od.sp_lo = a.limits[im];
			const uint32_t u1 = max(im, a.k2 ? (jp >> 1) : 0u) + (k - 1u - m);
			od.sp_hi = a.limits[min(u1, a.nl - 1u)];
		}
	}
	const uint32_t pos = first + lane;
	unsigned long long *pn = (k & 1u) ? f.pr0 : f.pr1;
	bool lbad = false;
	if (pos >= ws && pos < we && pos <= last) {
		const float old = od[pos];
		uint32_t cb = c * a.sb, S = a.sb;
		if (!a.sb) {   // both offsets requested together (one round trip)
			const uint32_t e0 = a.coff[c], e1 = a.coff[c + 1];
			cb = e0;
			S = e1 - e0;
		}
		const float best = relax_lane<false, 64, true, false, MS, G64>(a, (size_t) cb * 64 + lane, S, od, old);
		st_agent(pn + pos, ((unsigned long long) (k + 1u) << 32) | __float_as_uint(best));
		if (pos < wc) lbad = !(fabs((double) best - (double) old) / (double) old < a.tol);
	}
	if (__any(od.failed)) return FLOW_STOP;
	const uint32_t nbad = (uint32_t) __popcll(__ballot(lbad));
	if (lane == 0 && exact && wc > first) {
		const uint32_t tested = ((wc - 1u) >> 6) - (ws >> 6) + 1u;
		const unsigned long long o = atomicAdd(f.acc + k, (1ull << 32) | nbad);
		if ((uint32_t) (o >> 32) + 1u == tested) flow_publish(f, k, ik, ws, wc, we, jk, (uint32_t) o + nbad);
	}
	return FLOW_SWEPT;
}

// Entering sweep of a chunk: ring of its first position minus one (ring r
// enters the window in sweep r - 1), FLOW_NEVER past the rings.
__device__ __forceinline__ uint32_t flow_kin(const KArgs &a, uint32_t nr, uint32_t c)
{
	const uint32_t first = c << 6;
	if (first >= nr) return FLOW_NEVER;
	uint32_t lo = 0, hi = a.nl - 1u;   // limits[lo] <= first < limits[hi]
	while (hi - lo > 1u) {
		const uint32_t mid = (lo + hi) >> 1;
		if (a.limits[mid] <= first) lo = mid;
		else hi = mid;
	}
	return lo ? lo - 1u : 0u;
}

template <uint32_t MS, bool G64>
__global__ __launch_bounds__(64, 2) void k_flow(FlowArgs f)
{
	const KArgs &a = f.a;
	const uint32_t lane = threadIdx.x;
	const uint32_t w = blockIdx.x, nw = gridDim.x;
	const uint32_t nown = w < f.nchunks ? (f.nchunks - w + nw - 1u) / nw : 0u;   // <= FLOW_MAX_OWN
	const unsigned long long deadline = wall_clock64() + f.budget;
	// Lane q describes owned chunk q: its entering sweep and the entering
	// sweep of the chunk that ends the previous ring.
	uint32_t kin = FLOW_NEVER, pk = 0;
	if (lane < nown) {
		const uint32_t c = w + lane * nw;
		kin = flow_kin(a, f.nr, c);
		if (kin != FLOW_NEVER && kin > 0) {
			const uint32_t r = kin + 1u;   // ring of the first position (kin = r - 1 for r >= 1)
			pk = flow_kin(a, f.nr, (a.limits[r] - 1u) >> 6);
		}
	}
	uint32_t live = (uint32_t) __ballot(kin != FLOW_NEVER);
	uint32_t k = 0;
	while (live) {
		// Jump to the first sweep in which a live chunk is in the window.
		uint32_t kn = lane < 32u && ((live >> lane) & 1u) ? kin : FLOW_NEVER;
		for (uint32_t o = 32; o > 0; o >>= 1) kn = min(kn, (uint32_t) __shfl_xor((int) kn, (int) o));
		if (kn > k) k = kn;
		if (k >= a.max_iter) break;
		uint32_t todo = (uint32_t) __ballot(lane < 32u && ((live >> lane) & 1u) && kin <= k);
		uint32_t swept = 0;
		bool stop = false;
		while (todo) {
			const uint32_t q = (uint32_t) __builtin_ctz(todo);
			todo &= todo - 1u;
			const uint32_t kq = (uint32_t) __builtin_amdgcn_readlane((int) kin, (int) q);
			const uint32_t pq = (uint32_t) __builtin_amdgcn_readlane((int) pk, (int) q);
			const uint32_t r = flow_chunk<MS, G64>(f, w + q * nw, k, kq, pq, deadline);
			if (r == FLOW_OFF) live &= ~(1u << q);
			else if (r == FLOW_STOP) { stop = true; break; }
			else swept |= 1u << q;
		}
		if (stop) break;
		// The chunks' flags once their writes of sweep k have completed (one
		// wait per sweep, not per chunk: the readers of values inside the
		// window go by the tags; only frozen neighbours go by these flags);
		// entering chunks move the front.
		flow_wait_stores();
		if (lane < 32u && ((swept >> lane) & 1u)) {
			st_agent(f.cflag + (w + lane * nw), k + 1u);
			if (kin == k) atomicMax(&f.fs->front, k);
		}
		k++;
	}
}

// Pairs of both buffers from the initial distances (k_init, k_sources) of
// the n positions in rings, and the run's publication state.
__global__ void k_flow_init(const float *d0, unsigned long long *pr0, unsigned long long *pr1, uint32_t n,
                            uint32_t *cflag, uint32_t nchunks, uint32_t *hist, unsigned long long *acc, uint32_t nhist,
                            FlowState *fs)
{
	const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x, nt = gridDim.x * blockDim.x;
	for (uint32_t p = t; p < n; p += nt) {
		const unsigned long long v = __float_as_uint(d0[p]);
		pr0[p] = v;
		pr1[p] = v;
	}
	for (uint32_t c = t; c < nchunks; c += nt) cflag[c] = 0;
	for (uint32_t k = t; k < nhist; k += nt) {
		hist[k] = k == 0 ? 1u : 0u;   // state 0: i' = 1 (j = 2)
		acc[k] = 0;
	}
	if (t == 0) {
		fs->pub = 1;
		fs->abort = 0;
		fs->front = 0;
		fs->pad = 0;
	}
}

// Both float buffers of the n positions in rings from the pairs (the check
// sweeps and the result read them; later positions keep their INF).
__global__ void k_flow_unpack(const unsigned long long *pr0, const unsigned long long *pr1, float *d0, float *d1,
                              uint32_t n)
{
	for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
		d0[p] = __uint_as_float((uint32_t) pr0[p]);
		d1[p] = __uint_as_float((uint32_t) pr1[p]);
	}
}

}  // namespace gp
