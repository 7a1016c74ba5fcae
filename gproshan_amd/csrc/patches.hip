// Patch schedule: the relaxation of the PTP update run to its fixed point by
// spatially compact patches that iterate in LDS, with an active list.
//
// The vertices are renumbered in Morton order of their positions and cut
// into patches of P consecutive positions (one workgroup of P threads each).
// A launch processes the patches of the current active list: a patch loads
// its values into LDS, holds its out-of-patch neighbours at their values from
// global memory, and runs Jacobi iterations of the per-vertex update (the
// same FP32 solve with FP64 fallback as the ring sweeps) until no value moves
// by more than tol (relative) or local_iters iterations ran.  Values that
// dropped are written back in place (values only decrease, so a concurrent
// reader always sees a valid upper bound).  When a vertex dropped by more than
// tol, every patch holding one of its out-of-patch neighbours is appended to
// the next list; a patch that hit the iteration cap re-appends itself.  The
// launch that finds its list empty marks the solve done.
//
// Gating by a prediction (when the caller gives one): each patch carries a
// key, the smallest predicted distance dhat of its vertices.  A launch processes only the listed patches whose key
// is within band_width * h of the smallest key on the list and carries the
// others over, so the front moves through the mesh in predicted-distance
// order and fewer patches are recomputed after a later, shorter arrival.
// A prediction only reorders work: the fixed point reached is the same.
//
// Lists rotate over four slots (launch k reads list k&3, appends to
// (k+1)&3, clears the count and key minimum of (k+2)&3, which launch k-2
// consumed), so the launches can be replayed from graphs.  The
// reference has no counterpart: it runs the ring window schedule only
// (src/cuda/geodesics_ptp.cu:198-231); results are compared with the fixed
// point of the same relaxation.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <numeric>
#include <vector>

#include "device.hpp"
#include "sweep_dev.hpp"

namespace gp {

struct PatchState {
	uint32_t count[4];
	uint32_t kmin[4];       // float bits of the smallest key on each list
	uint32_t done, launches, pad0, pad1;
	unsigned long long visits, updates, deferred, pad2;
};

struct PArgs {
	const uint32_t *coff, *cid;
	const uint16_t *lidx;
	const float2 *cgeo;
	const double4 *cgeo64;
	const Edge64 *cx64;    // unused (FP32 only); keeps relax helpers' interface
	const float *pkey;
	float *dist;
	uint32_t *list;        // [4][npatch]
	uint32_t *mark;        // [4][npatch]
	uint8_t *defer;        // [npatch] launches a listed patch has waited
	PatchState *st;
	uint32_t *trace;       // [max_launch][8]
	float h, inv_h;
	float tol, width;      // width = band_width * h, +inf without gating
	uint32_t n, npatch, local_iters, max_launch;
};

__device__ __forceinline__ void p_append(const PArgs &a, uint32_t slot, uint32_t q)
{
	if (atomicExch(&a.mark[(size_t) slot * a.npatch + q], 1u) == 0u) {
		const uint32_t at = atomicAdd(&a.st->count[slot], 1u);
		a.list[(size_t) slot * a.npatch + at] = q;
		atomicMin(&a.st->kmin[slot], __float_as_uint(a.pkey[q]));
	}
}

// Per-slot word of a patch visit: bits 0-15 the neighbour's index in the
// patch (LOCAL_OUT if outside), bit 16 the slot is empty, bit 17 triangle
// (k, k+1) takes the FP64 update.
constexpr uint32_t PK_EMPTY = 1u << 16, PK_F64 = 1u << 17;
constexpr uint32_t FCACHE = 3;     // FP64 triangles per vertex held in LDS
constexpr uint32_t WAKE_TAB = 64;  // LDS set of the patches a visit wakes
constexpr uint32_t MAX_DEFER = 8;  // a listed patch waits at most this many launches

template <uint32_t P>
__global__ __launch_bounds__(P) void k_patch_sweep(PArgs a, uint32_t k4)
{
	__shared__ float sd[2][P];
	// FP64 geometry of each thread's first FCACHE FP64-flagged triangles,
	// loaded once per visit, and their inputs of the current iteration (any
	// further flagged triangle of a vertex -- rare -- reads global memory).
	__shared__ double fq0[FCACHE][P], fq1[FCACHE][P], fg[FCACHE][P], fdet[FCACHE][P], fS[FCACHE][P];
	__shared__ float fta[FCACHE][P], ftb[FCACHE][P];
	__shared__ uint32_t s_wake[WAKE_TAB];
	__shared__ uint8_t chg[2][P];        // value dropped in the iteration that wrote sd[b]
	__shared__ uint32_t s_cap, s_entry[2], s_upd;
	PatchState *st = a.st;
	const uint32_t cur = k4, nxt = (k4 + 1u) & 3u, rst = (k4 + 2u) & 3u;
	const uint32_t count = st->count[cur];
	const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
	if (count == 0) {
		if (lead) st->done = 1;
		return;
	}
	const float T = __uint_as_float(st->kmin[cur]) + a.width;
	if (lead) {
		const uint32_t l = st->launches;
		if (a.trace && l < a.max_launch) {
			uint32_t *r = a.trace + 8 * (size_t) l;
			r[0] = l + 1; r[1] = count; r[2] = st->kmin[cur]; r[3] = __float_as_uint(T);
			r[4] = 0; r[5] = 0; r[6] = 0; r[7] = 0;
		}
		st->launches = l + 1;
		st->count[rst] = 0;
		st->kmin[rst] = 0x7F800000u;
	}
	const uint32_t tid = threadIdx.x, lane = tid & 63u;
	const float h = a.h, inv_h = a.inv_h;
	for (uint32_t idx = blockIdx.x; idx < count; idx += gridDim.x) {
		// Thread 0 takes the entry and decides for the workgroup: a patch
		// outside the band is carried over (at most MAX_DEFER times running).
		if (tid == 0) {
			const uint32_t q = a.list[(size_t) cur * a.npatch + idx];
			a.mark[(size_t) cur * a.npatch + q] = 0;
			const bool go = a.pkey[q] <= T || a.defer[q] >= MAX_DEFER;
			if (!go) {
				a.defer[q]++;
				p_append(a, nxt, q);
				atomicAdd(&st->deferred, 1ull);
			}
			s_entry[0] = q;
			s_entry[1] = go ? 1u : 0u;
		}
		__syncthreads();
		const uint32_t p = s_entry[0];
		const bool go = s_entry[1] != 0;
		__syncthreads();
		if (!go) continue;
		const uint32_t pos = p * P + tid;
		const bool live = pos < a.n;
		uint32_t S = 0;
		size_t base = 0;
		if (live) {
			const uint32_t c = pos >> 6, cb = a.coff[c];
			S = a.coff[c + 1] - cb;
			base = (size_t) cb * 64 + lane;
		}
		const float v0 = live ? a.dist[pos] : (float) INFINITY;
		const bool fast = S <= MAXS;
		uint32_t pk[MAXS];
		float2 g[MAXS];
		float tout[MAXS];
		uint32_t fks = 0xFFFu, nf = 0;       // cached FP64 slots, 4 bits each (0xF: none)
#pragma unroll
		for (uint32_t k = 0; k < MAXS; k++) {
			pk[k] = LOCAL_OUT | PK_EMPTY;
			tout[k] = (float) INFINITY;
			if (fast && k < S) {
				const uint32_t id = a.cid[base + 64 * k];
				const uint32_t li = a.lidx[base + 64 * k];
				pk[k] = li | ((id & SLOT_EMPTY) ? PK_EMPTY : 0u) | ((id & F64_FLAG) ? PK_F64 : 0u);
				g[k] = a.cgeo[base + 64 * k];
				if (li == LOCAL_OUT && !(id & SLOT_EMPTY)) tout[k] = a.dist[id & SLOT_ID_MASK];
				if ((id & F64_FLAG) && nf < FCACHE) {
					fks = (fks & ~(0xFu << (4 * nf))) | (k << (4 * nf));
					nf++;
				}
			}
		}
#pragma unroll 1
		for (uint32_t c = 0; c < nf; c++) {
			const size_t at = base + 64 * (size_t) ((fks >> (4 * c)) & 0xFu);
			const double4 gd = a.cgeo64[at];
			fq0[c][tid] = gd.x; fg[c][tid] = gd.y; fdet[c][tid] = gd.z; fS[c][tid] = gd.w;
			fq1[c][tid] = a.cgeo64[at + 64].x;
		}
		sd[0][tid] = v0;
		if (tid < WAKE_TAB) s_wake[tid] = 0;
		if (tid == 0) { s_cap = 0; s_upd = 0; }
		__syncthreads();
		uint32_t b = 0, it = 0;
		for (; it < a.local_iters; it++) {
			const float curv = sd[b][tid];
			float best = curv;
			// A vertex is recomputed only when an in-patch neighbour dropped in
			// the last iteration (out-of-patch values are fixed during the
			// visit): otherwise its inputs are those of its last evaluation,
			// whose result it already holds, so the skip is exact.
			bool need = it == 0 || !fast;
			if (!need) {
#pragma unroll
				for (uint32_t k = 0; k < MAXS; k++) {
					const uint32_t lk = pk[k] & 0xFFFFu;
					if (k < S && lk != LOCAL_OUT && chg[b][lk]) need = true;
				}
			}
			{
				const unsigned long long bal = __ballot(need && live);
				if (lane == 0 && bal) atomicAdd(&s_upd, (uint32_t) __popcll(bal));
			}
			if (need && fast) {
				// FP32 triangles inline; the FP64-flagged ones after the slot
				// loop, so the FP64 code is emitted once, not per unrolled slot.
				float ta = (pk[0] & 0xFFFFu) != LOCAL_OUT ? sd[b][pk[0] & 0xFFFFu] : tout[0];
				uint32_t fmore = 0;
#pragma unroll
				for (uint32_t k = 0; k + 1 < MAXS; k++) {
					if (k + 1 < S && !(pk[k + 1] & PK_EMPTY)) {
						const uint32_t lb = pk[k + 1] & 0xFFFFu;
						const float tb = lb != LOCAL_OUT ? sd[b][lb] : tout[k + 1];
						if (!(pk[k] & PK_F64)) {
							const float q = update_stable_q(g[k].x, g[k + 1].x, g[k].y, ta, tb, h, inv_h);
							if (q < best) best = q;
						} else if ((fks & 0xFu) == k) {
							fta[0][tid] = ta; ftb[0][tid] = tb;
						} else if (((fks >> 4) & 0xFu) == k) {
							fta[1][tid] = ta; ftb[1][tid] = tb;
						} else if (((fks >> 8) & 0xFu) == k) {
							fta[2][tid] = ta; ftb[2][tid] = tb;
						} else {
							fmore |= 1u << k;
						}
						ta = tb;
					}
				}
#pragma unroll 1
				for (uint32_t c = 0; c < nf; c++) {
					const float q = (float) update_stable<double>(fq0[c][tid], fq1[c][tid], fg[c][tid], fdet[c][tid],
					                                              fS[c][tid], (double) fta[c][tid], (double) ftb[c][tid],
					                                              (double) h, (double) inv_h);
					if (q < best) best = q;
				}
				while (fmore) {
					const uint32_t k = __builtin_ctz(fmore);
					fmore &= fmore - 1u;
					const size_t at = base + 64 * (size_t) k;
					const uint32_t la = a.lidx[at], lb = a.lidx[at + 64];
					const float tka = la != LOCAL_OUT ? sd[b][la] : a.dist[a.cid[at] & SLOT_ID_MASK];
					const float tkb = lb != LOCAL_OUT ? sd[b][lb] : a.dist[a.cid[at + 64] & SLOT_ID_MASK];
					const float q = tri_flagged(a, at, at + 64, tka, tkb);
					if (q < best) best = q;
				}
			} else if (!fast) {
				// Fans longer than MAXS: stream the slots.
				uint32_t ida = a.cid[base];
				uint32_t lia = a.lidx[base];
				float2 ga = a.cgeo[base];
				float ta = lia != LOCAL_OUT ? sd[b][lia] : a.dist[ida & SLOT_ID_MASK];
				for (uint32_t k = 1; k < S; k++) {
					const uint32_t idb = a.cid[base + 64 * k];
					if (idb & SLOT_EMPTY) break;
					const uint32_t lib = a.lidx[base + 64 * k];
					const float2 gb = a.cgeo[base + 64 * k];
					const float tb = lib != LOCAL_OUT ? sd[b][lib] : a.dist[idb & SLOT_ID_MASK];
					const float q = (ida & F64_FLAG) ? tri_flagged(a, base + 64 * (k - 1), base + 64 * k, ta, tb)
					                                 : update_stable_q(ga.x, gb.x, ga.y, ta, tb, h, inv_h);
					if (q < best) best = q;
					ida = idb; ga = gb; ta = tb;
				}
			}
			sd[b ^ 1u][tid] = best;
			chg[b ^ 1u][tid] = best < curv ? 1 : 0;
			const int moved = __syncthreads_or(best < curv && (curv - best) > a.tol * best);
			b ^= 1u;
			if (!moved) break;
		}
		const float v = sd[b][tid];
		if (live && v < v0) a.dist[pos] = v;
		// Patches across the boundary of every vertex that moved: collected
		// into an LDS set first, so each is appended once per visit.
		if (live && v < v0 && (!(v0 < INFINITY) || (v0 - v) > a.tol * v)) {
			for (uint32_t k = 0; k < S; k++) {
				const uint32_t idk = a.cid[base + 64 * k];
				if ((idk & SLOT_EMPTY) || a.lidx[base + 64 * k] != LOCAL_OUT) continue;
				const uint32_t q = (idk & SLOT_ID_MASK) / P;
				uint32_t hsh = q & (WAKE_TAB - 1u), probe = 0;
				for (; probe < WAKE_TAB; probe++) {
					const uint32_t old = atomicCAS(&s_wake[hsh], 0u, q + 1u);
					if (old == 0u || old == q + 1u) break;
					hsh = (hsh + 1u) & (WAKE_TAB - 1u);
				}
				if (probe == WAKE_TAB) p_append(a, nxt, q);   // table full: append directly
			}
		}
		if (it == a.local_iters && tid == 0) s_cap = 1;
		__syncthreads();
		if (tid < WAKE_TAB && s_wake[tid]) p_append(a, nxt, s_wake[tid] - 1u);
		if (tid == 0) {
			a.defer[p] = 0;
			if (s_cap) p_append(a, nxt, p);
			atomicAdd(&st->visits, 1ull);
			atomicAdd(&st->updates, (unsigned long long) s_upd);
		}
	}
}

__global__ void k_patch_init(float *dist, uint32_t n, uint32_t *mark, uint32_t nmark, uint8_t *defer, uint32_t np,
                             PatchState *st)
{
	const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
	for (uint32_t v = t; v < n; v += gridDim.x * blockDim.x) dist[v] = INFINITY;
	for (uint32_t q = t; q < nmark; q += gridDim.x * blockDim.x) mark[q] = 0;
	for (uint32_t q = t; q < np; q += gridDim.x * blockDim.x) defer[q] = 0;
	if (t == 0) {
		uint32_t *w = (uint32_t *) st;
		for (uint32_t q = 0; q < sizeof(PatchState) / 4; q++) w[q] = 0;
		for (uint32_t q = 0; q < 4; q++) st->kmin[q] = 0x7F800000u;
	}
}

// Sources get 0 and their patches form list 0 (one thread: ns is small).
__global__ void k_patch_seed(PArgs a, const uint32_t *src, uint32_t ns, uint32_t P)
{
	if (threadIdx.x != 0 || blockIdx.x != 0) return;
	for (uint32_t q = 0; q < ns; q++) {
		a.dist[src[q]] = 0.f;
		const uint32_t p = src[q] / P;
		if (a.mark[p] == 0) {
			a.mark[p] = 1;
			a.list[a.st->count[0]++] = p;
			a.st->kmin[0] = min(a.st->kmin[0], __float_as_uint(a.pkey[p]));
		}
	}
}

__global__ void k_patch_out(const float *dist, double *out, uint32_t n)
{
	const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
	for (uint32_t v = t; v < n; v += gridDim.x * blockDim.x) out[v] = (double) dist[v];
}

// ---------------------------------------------------------------------------

struct PatchPlan {
	Che *mesh = nullptr;
	std::unique_ptr<Che> ren;          // Morton-numbered mesh
	std::vector<uint32_t> inv;         // original -> Morton numbering
	uint32_t n = 0, nv_orig = 0, ns = 0, P = 256, npatch = 0, grid = 0, max_launch = 0;
	gp_ptp_opts opts{};
	float width = INFINITY;
	uint32_t *d_coff = nullptr, *d_cid = nullptr, *d_list = nullptr, *d_mark = nullptr, *d_trace = nullptr,
	         *d_src = nullptr;
	uint16_t *d_lidx = nullptr;
	uint8_t *d_defer = nullptr;
	float2 *d_cgeo = nullptr;
	double4 *d_cgeo64 = nullptr;
	float *d_pkey = nullptr, *d_dist = nullptr;
	double *d_out = nullptr;
	PatchState *d_state = nullptr, *h_state = nullptr;
	hipStream_t stream = nullptr;
	hipEvent_t e0 = nullptr, e1 = nullptr;
	hipGraphExec_t graph = nullptr;
	bool has_run = false;

	~PatchPlan()
	{
		for (void *p : { (void *) d_coff, (void *) d_cid, (void *) d_list, (void *) d_mark, (void *) d_trace,
		                 (void *) d_src, (void *) d_lidx, (void *) d_cgeo, (void *) d_cgeo64, (void *) d_pkey,
		                 (void *) d_dist, (void *) d_out, (void *) d_state, (void *) d_defer })
			(void) hipFree(p);
		if (h_state) (void) hipHostFree(h_state);
		if (graph) (void) hipGraphExecDestroy(graph);
		if (e0) (void) hipEventDestroy(e0);
		if (e1) (void) hipEventDestroy(e1);
		if (stream) (void) hipStreamDestroy(stream);
	}
};

template <typename V>
static int pupload(V *&dst, const V *src, size_t n)
{
	if (hipMalloc(&dst, sizeof(V) * std::max<size_t>(n, 1)) != hipSuccess) return fail(GP_ERR_NOMEM, "patch plan: allocation failed");
	if (n) GP_HIP(hipMemcpy(dst, src, sizeof(V) * n, hipMemcpyHostToDevice));
	return GP_OK;
}

template <uint32_t P>
static int occupancy(int &occ)
{
	GP_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_patch_sweep<P>, P, 0));
	return GP_OK;
}

int patch_create(Che *mesh, const uint32_t *src, uint32_t ns, const double *dhat, const gp_ptp_opts &o,
                 PatchPlan **out)
{
	*out = nullptr;
	if (!mesh || !src || !ns) return fail(GP_ERR_INVALID, "patch plan: no sources");
	if (o.precision != GP_PREC_F32) return fail(GP_ERR_INVALID, "patch and band schedules run in FP32");
	if (o.local_iters == 0 || o.local_iters > 4096) return fail(GP_ERR_INVALID, "local_iters out of range");
	if (!(o.tol > 0) || !(o.tol < 1)) return fail(GP_ERR_INVALID, "tol must be in (0, 1)");
	const uint32_t P = o.block ? o.block : 256;
	if (P != 256 && P != 512) return fail(GP_ERR_INVALID, "patch size (block) must be 256 or 512");
	std::vector<uint8_t> seen(mesh->nv, 0);
	for (uint32_t q = 0; q < ns; q++) {
		if (src[q] >= mesh->nv) return fail(GP_ERR_INVALID, "patch plan: source %u out of range", src[q]);
		if (seen[src[q]]) return fail(GP_ERR_DUPLICATE_SOURCE, "patch plan: source %u listed twice", src[q]);
		seen[src[q]] = 1;
	}
	auto pp = std::make_unique<PatchPlan>();
	pp->mesh = mesh;
	pp->opts = o;
	pp->nv_orig = mesh->nv;
	pp->ns = ns;
	pp->P = P;
	std::vector<uint32_t> order;
	morton_order(*mesh, order);
	pp->ren = std::make_unique<Che>();
	int rc = coalesce(*mesh, order.data(), mesh->nv, *pp->ren, pp->inv);
	if (rc) return rc;
	pp->ren->mean_edge = mesh->mean_edge;   // the same h as the mesh itself
	Che &m = *pp->ren;
	pp->n = m.nv;
	rc = upload_mesh(m);
	if (rc) return rc;
	Layout L;
	if ((rc = build_layout(m, nullptr, m.nv, false, L, P))) return rc;
	pp->npatch = (m.nv + P - 1) / P;
	// Patch keys: smallest predicted distance of the patch (0 without gating).
	std::vector<float> pkey(pp->npatch, 0.f);
	if (dhat) {
		std::fill(pkey.begin(), pkey.end(), INFINITY);
		for (uint32_t v = 0; v < mesh->nv; v++) {
			const double d = dhat[v];
			const float f = std::isfinite(d) && d > 0 ? (float) d : (std::isnan(d) ? INFINITY : (float) std::max(d, 0.0));
			float &k = pkey[pp->inv[v] / P];
			k = std::min(k, f);
		}
		for (uint32_t q = 0; q < ns; q++) pkey[pp->inv[src[q]] / P] = 0.f;
		if (!(o.band_width > 0)) return fail(GP_ERR_INVALID, "band_width must be positive");
		pp->width = (float) (o.band_width * m.mean_edge);
	}
	std::vector<uint32_t> srcm(ns);
	for (uint32_t q = 0; q < ns; q++) srcm[q] = pp->inv[src[q]];
	int dev = 0, ncu = 0, occ = 0;
	GP_HIP(hipGetDevice(&dev));
	GP_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
	rc = P == 256 ? occupancy<256>(occ) : occupancy<512>(occ);
	if (rc) return rc;
	pp->grid = o.grid ? o.grid : (uint32_t) ncu * (uint32_t) std::max(1, occ);
	pp->max_launch = o.max_sweeps ? o.max_sweeps : 64 * pp->npatch + 1024;
	std::vector<float2> cg(L.cgeo.size());
	std::memcpy(cg.data(), L.cgeo.data(), sizeof(float2) * cg.size());
	std::vector<double4> cg64(L.cgeo64.size());
	std::memcpy(cg64.data(), L.cgeo64.data(), sizeof(double4) * cg64.size());
	if ((rc = pupload(pp->d_coff, L.coff.data(), L.coff.size())) || (rc = pupload(pp->d_cid, L.cid.data(), L.cid.size())) ||
	    (rc = pupload(pp->d_lidx, L.lidx.data(), L.lidx.size())) || (rc = pupload(pp->d_cgeo, cg.data(), cg.size())) ||
	    (rc = pupload(pp->d_cgeo64, cg64.data(), cg64.size())) || (rc = pupload(pp->d_pkey, pkey.data(), pkey.size())) ||
	    (rc = pupload(pp->d_src, srcm.data(), srcm.size())))
		return rc;
	const uint32_t trace_rows = std::min<uint32_t>(pp->max_launch, 1u << 20);
	if (hipMalloc(&pp->d_list, 4ull * 4 * pp->npatch) != hipSuccess ||
	    hipMalloc(&pp->d_mark, 4ull * 4 * pp->npatch) != hipSuccess || hipMalloc(&pp->d_defer, pp->npatch) != hipSuccess ||
	    hipMalloc(&pp->d_trace, 32ull * trace_rows) != hipSuccess || hipMalloc(&pp->d_dist, 4ull * pp->n) != hipSuccess ||
	    hipMalloc(&pp->d_out, 8ull * pp->n) != hipSuccess || hipMalloc(&pp->d_state, sizeof(PatchState)) != hipSuccess ||
	    hipHostMalloc(&pp->h_state, sizeof(PatchState)) != hipSuccess)
		return fail(GP_ERR_NOMEM, "patch plan: device allocation failed");
	pp->max_launch = std::min(pp->max_launch, trace_rows);
	GP_HIP(hipStreamCreateWithFlags(&pp->stream, hipStreamNonBlocking));
	GP_HIP(hipEventCreate(&pp->e0));
	GP_HIP(hipEventCreate(&pp->e1));
	*out = pp.release();
	return GP_OK;
}

static PArgs patch_args(const PatchPlan *pp)
{
	PArgs a{};
	a.coff = pp->d_coff; a.cid = pp->d_cid; a.lidx = pp->d_lidx; a.cgeo = pp->d_cgeo; a.cgeo64 = pp->d_cgeo64;
	a.cx64 = nullptr; a.pkey = pp->d_pkey; a.dist = pp->d_dist; a.list = pp->d_list; a.mark = pp->d_mark;
	a.defer = pp->d_defer;
	a.st = pp->d_state; a.trace = pp->d_trace;
	a.h = pp->ren->dev->h; a.inv_h = pp->ren->dev->inv_h;
	a.tol = (float) pp->opts.tol; a.width = pp->width;
	a.n = pp->n; a.npatch = pp->npatch; a.local_iters = pp->opts.local_iters; a.max_launch = pp->max_launch;
	return a;
}

constexpr uint32_t PATCH_GRAPH = 32;  // launches per graph replay (a multiple of 4)

static void enqueue_patch(const PatchPlan *pp, const PArgs &a, uint32_t count)
{
	for (uint32_t q = 0; q < count; q++) {
		const uint32_t k4 = q & 3u;
		if (pp->P == 256) k_patch_sweep<256><<<pp->grid, 256, 0, pp->stream>>>(a, k4);
		else k_patch_sweep<512><<<pp->grid, 512, 0, pp->stream>>>(a, k4);
	}
}

int patch_run(PatchPlan *pp, gp_ptp_stats *stats)
{
	if (!pp) return fail(GP_ERR_INVALID, "plan is null");
	const auto t0 = std::chrono::steady_clock::now();
	const PArgs a = patch_args(pp);
	const bool graph = pp->opts.launch_mode == GP_LAUNCH_GRAPH && graphs_enabled();
	if (graph && !pp->graph) {
		PArgs pa = a;
		const uint32_t k4s[4] = { 0, 1, 2, 3 };
		const void *fn = pp->P == 256 ? (const void *) k_patch_sweep<256> : (const void *) k_patch_sweep<512>;
		GraphChain gc;
		for (uint32_t q = 0; q < PATCH_GRAPH; q++) {
			void *args[] = { &pa, (void *) &k4s[q & 3u] };
			gc.add(fn, dim3(pp->grid), dim3(pp->P), args);
		}
		const hipError_t e = gc.instantiate(&pp->graph);
		if (e != hipSuccess) return fail(GP_ERR_HIP, "patch graph: %s", hipGetErrorString(e));
	}
	const uint32_t ig = std::max<uint32_t>(1, std::min<uint32_t>(1024, (pp->n + 255) / 256));
	k_patch_init<<<ig, 256, 0, pp->stream>>>(pp->d_dist, pp->n, pp->d_mark, 4 * pp->npatch, pp->d_defer, pp->npatch,
	                                         pp->d_state);
	k_patch_seed<<<1, 64, 0, pp->stream>>>(a, pp->d_src, pp->ns, pp->P);
	GP_HIP(hipGetLastError());
	GP_HIP(hipEventRecord(pp->e0, pp->stream));
	uint32_t launched = 0;
	for (;;) {
		if (graph) GP_HIP(hipGraphLaunch(pp->graph, pp->stream));
		else enqueue_patch(pp, a, PATCH_GRAPH);
		launched += PATCH_GRAPH;
		GP_HIP(hipGetLastError());
		snapshot(pp->d_state, pp->h_state, pp->stream);
		GP_HIP(hipGetLastError());
		GP_HIP(hipStreamSynchronize(pp->stream));
		if (pp->h_state->done) break;
		if (launched >= pp->max_launch)
			return fail(GP_ERR_KERNEL, "patch schedule did not finish in %u launches", launched);
	}
	GP_HIP(hipEventRecord(pp->e1, pp->stream));
	GP_HIP(hipStreamSynchronize(pp->stream));
	float ms = 0;
	GP_HIP(hipEventElapsedTime(&ms, pp->e0, pp->e1));
	pp->has_run = true;
	if (stats) {
		const PatchState &s = *pp->h_state;
		stats->sweeps = s.launches;
		stats->n_limits = pp->npatch;
		stats->relaxations = s.updates;
		stats->cond_checks = s.visits * pp->P;
		stats->f64_updates = 0;
		stats->patch_visits = s.visits;
		stats->deferred = s.deferred;
		stats->kernel_ms = ms;
		stats->total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
		stats->precision = GP_PREC_F32;
		stats->capped = 0;
		stats->kernel = 0;
		stats->check_sweeps = 0;
	}
	return GP_OK;
}

int patch_result(PatchPlan *pp, double *dist)
{
	if (!pp || !dist) return fail(GP_ERR_INVALID, "plan_result: null argument");
	if (!pp->has_run) return fail(GP_ERR_INVALID, "plan_result: plan has not run");
	const uint32_t g = std::max<uint32_t>(1, std::min<uint32_t>(1024, (pp->n + 255) / 256));
	k_patch_out<<<g, 256, 0, pp->stream>>>(pp->d_dist, pp->d_out, pp->n);
	GP_HIP(hipGetLastError());
	std::vector<double> tmp(pp->n);
	GP_HIP(hipMemcpyAsync(tmp.data(), pp->d_out, sizeof(double) * pp->n, hipMemcpyDeviceToHost, pp->stream));
	GP_HIP(hipStreamSynchronize(pp->stream));
	for (uint32_t v = 0; v < pp->nv_orig; v++) dist[v] = tmp[pp->inv[v]];
	return GP_OK;
}

int patch_trace(PatchPlan *pp, uint32_t *out, uint32_t cap, uint32_t *n)
{
	if (!pp || !pp->has_run) return fail(GP_ERR_INVALID, "plan_trace: plan has not run");
	const uint32_t s = std::min(pp->h_state->launches, pp->max_launch);
	if (n) *n = s;
	const uint32_t c = std::min(s, cap);
	if (c && out) {
		GP_HIP(hipMemcpyAsync(out, pp->d_trace, 32ull * c, hipMemcpyDeviceToHost, pp->stream));
		GP_HIP(hipStreamSynchronize(pp->stream));
	}
	return GP_OK;
}

void patch_destroy(PatchPlan *pp) { delete pp; }

}  // namespace gp
