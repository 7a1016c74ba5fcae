// Batched PTP: many (mesh, source) fields of uneven size in one persistent
// launch (a hipGraph of its queue reset and the kernel).
//
// Each 1024-thread workgroup repeatedly takes the next job from an atomic
// queue (jobs sorted by decreasing cost V^1.5, so the grid drains evenly), and
// runs it entirely by itself: the ring build (the same BFS as the toplesets
// kernel), then the reference window schedule (K/2 band limit, PTP_TOL test,
// 2*|limits| cap, last-written buffer) with __syncthreads as the only barrier,
// on the rings of the reference's non-coalesced driver
// (src/cuda/geodesics_ptp.cu:174-299).  Each mesh is stored vertex-major (a
// vertex's slots are contiguous) in Morton order; after its BFS the workgroup
// gathers the field's own ring-ordered tables into its scratch (64 positions
// to a chunk, one lane per vertex, neighbour ids mapped to positions, as the
// coalesced driver renumbers the mesh, src/geodesics_ptp.cpp:15-38), so every
// slot load of a sweep is one contiguous access; the FP64 geometry of flagged
// triangles is read from the mesh's tables through the slot's index there.
// The gather keeps every fan's order, hence the rings and the arithmetic, and
// the output is scattered back.  Every workgroup leaves the loop when the
// queue is exhausted, so the launch always drains.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <numeric>
#include <vector>

#include "device.hpp"
#include "rings_dev.hpp"
#include "sweep_dev.hpp"

namespace gp {

struct BMesh {
	const uint32_t *fan_off, *fan;
	const uint8_t *closed;
	const uint32_t *coff, *cid;
	const float2 *cgeo;
	const double4 *cgeo64;
	const Edge64 *cx64;
	const uint32_t *perm;   // Morton position -> original vertex id
	float h, inv_h;
	uint32_t nv, ms;        // ms: the longest fan (slots of a field's chunk)
};

struct BJob {
	uint32_t mesh, source, sweeps, pad;
	unsigned long long out, relax;
};

struct BArgs {
	const BMesh *meshes;
	BJob *jobs;
	uint32_t njobs, max_nv;
	uint32_t *queue;
	uint32_t *level, *sorted, *limits, *nlim;
	unsigned long long *key;
	float *dist;
	float *out;
	double tol;
	// Per-workgroup field tables (fslots slots each): ids in positions, FP32
	// geometry, the slot's index in the mesh's tables; inv: vertex -> position.
	uint32_t *fcid, *fmidx, *finv;
	float2 *fgeo;
	size_t fslots;
};

// The field's tables as relax_lane reads them.
struct FieldTables {
	const uint32_t *cid;
	const float2 *cgeo;
	GeoRemap64 cgeo64;
	const Edge64 *cx64;
	float h, inv_h;
};

__global__ void k_batch_reset(uint32_t *queue) { *queue = 0; }

// NT threads per field (a multiple of 64, at most 1024).
template <uint32_t NT, uint32_t MS>
__global__ __launch_bounds__(NT) void k_batch(BArgs a)
{
	__shared__ uint32_t s_job, s_bad;
	const uint32_t tid = threadIdx.x, nt = NT, wg = blockIdx.x;
	const size_t mv = a.max_nv;
	uint32_t *level = a.level + wg * mv;
	uint32_t *sorted = a.sorted + wg * mv;
	uint32_t *limits = a.limits + wg * (mv + 2);
	unsigned long long *key = a.key + wg * mv;
	float *d0 = a.dist + 2 * wg * mv, *d1 = d0 + mv;
	uint32_t *fcid = a.fcid + wg * a.fslots, *fmidx = a.fmidx + wg * a.fslots, *finv = a.finv + wg * mv;
	float2 *fgeo = a.fgeo + wg * a.fslots;
	// Values every thread of the workgroup holds alike (the wave index, the
	// job, the ring-test count) are taken from the first lane, so the
	// compiler keeps them and what follows from them (the mesh, the window,
	// the chunk index, the slot count) in scalar registers.
	const uint32_t lane = tid & 63u, wave = __builtin_amdgcn_readfirstlane(tid >> 6), nwave = NT >> 6;
	for (;;) {
		if (tid == 0) s_job = atomicAdd(a.queue, 1u);
		__syncthreads();
		const uint32_t job = __builtin_amdgcn_readfirstlane(s_job);
		__syncthreads();
		if (job >= a.njobs) break;
		const uint32_t src = a.jobs[job].source;
		const BMesh M = a.meshes[a.jobs[job].mesh];
		const uint32_t nv = M.nv;
		for (uint32_t v = tid; v < nv; v += nt) {
			level[v] = GP_NIL; key[v] = ~0ull; sorted[v] = GP_NIL; finv[v] = GP_NIL;
			d0[v] = INFINITY; d1[v] = INFINITY;
		}
		__syncthreads();
		const BfsJob bj{ &a.jobs[job].source, 1u, level, key, sorted, limits, a.nlim + wg };
		bfs_block(M.fan_off, M.fan, M.closed, bj);
		const uint32_t nl = __hip_atomic_load(a.nlim + wg, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
		// The field's ring-ordered tables: position p holds vertex sorted[p]
		// (the source is position 0).
		const uint32_t nr = nl ? limits[nl - 1] : 0u;
		for (uint32_t q = tid; q < nr; q += nt) finv[sorted[q]] = q;
		__syncthreads();
		const uint32_t S = M.ms, npos = (nr + 63u) & ~63u;
		for (uint32_t q = tid; q < npos; q += nt) {
			const size_t fb = (size_t) (q >> 6) * S * 64 + (q & 63u);
			uint32_t mb = 0, n = 0;
			if (q < nr) {
				const uint32_t v = sorted[q];
				mb = M.coff[v];
				n = M.coff[v + 1] - mb;
			}
			for (uint32_t k = 0; k < S; k++) {
				const size_t at = fb + 64 * (size_t) k;
				if (k < n) {
					const uint32_t id = M.cid[mb + k];
					fcid[at] = finv[id & SLOT_ID_MASK] | (id & ~SLOT_ID_MASK);
					fgeo[at] = M.cgeo[mb + k];
					fmidx[at] = mb + k;
				} else {
					fcid[at] = q | SLOT_EMPTY;
					fgeo[at] = make_float2(0.f, 0.f);
					fmidx[at] = 0;
				}
			}
		}
		if (tid == 0 && nr) { d0[0] = 0.f; d1[0] = 0.f; }
		__syncthreads();
		const FieldTables F{ fcid, fgeo, GeoRemap64{ M.cgeo64, fmidx }, nullptr, M.h, M.inv_h };
		uint32_t i = 1, j = 2, d = 0, iter = 0;
		unsigned long long relax = 0;
		const uint32_t max_iter = nl << 1;
		bool done = nl < 3;
		while (!done) {
			const uint32_t ii = i < (j >> 1) ? (j >> 1) : i;
			const uint32_t start = limits[ii], end = limits[j], cend = limits[ii + 1];
			const float *od = d ? d1 : d0;
			float *nd = d ? d0 : d1;
			if (tid == 0) s_bad = 0;
			__syncthreads();
			uint32_t bad = 0;
			for (uint32_t c = (start >> 6) + wave; c < ((end + 63u) >> 6); c += nwave) {
				const uint32_t pos = (c << 6) + lane;
				bool lbad = false;
				if (pos >= start && pos < end) {
					const float old = od[pos];
					const float best = relax_lane<false, 64, false, false, MS>(F, (size_t) c * S * 64 + lane, S, od, old);
					nd[pos] = best;
					if (pos < cend) lbad = !(fabs((double) best - (double) old) / (double) old < a.tol);
				}
				bad += (uint32_t) __popcll(__ballot(lbad));
			}
			if (lane == 0 && bad) atomicAdd(&s_bad, bad);
			__syncthreads();
			const uint32_t nbad = __builtin_amdgcn_readfirstlane(s_bad);
			relax += end - start;
			i = ii + (nbad == 0 ? 1u : 0u);
			if (j < nl - 1) j++;
			d ^= 1u;
			iter++;
			done = !(i < j) || !(iter < max_iter);
			__syncthreads();
		}
		const float *res = d ? d1 : d0;            // the buffer the last sweep wrote
		float *out = a.out + a.jobs[job].out;
		for (uint32_t v = tid; v < nv; v += nt) out[M.perm[v]] = finv[v] != GP_NIL ? res[finv[v]] : INFINITY;
		if (tid == 0) { a.jobs[job].sweeps = iter; a.jobs[job].relax = relax; }
		__syncthreads();
	}
}

}  // namespace gp

// ---------------------------------------------------------------------------

struct gp_batch {
	struct MeshTables {
		std::unique_ptr<gp::Che> ren;   // Morton-numbered copy of the mesh
		std::vector<uint32_t> inv;      // original -> Morton numbering
		uint32_t *coff = nullptr, *cid = nullptr, *perm = nullptr;
		float2 *cgeo = nullptr;
		double4 *cgeo64 = nullptr;
	};
	std::vector<MeshTables> tabs;
	std::vector<unsigned long long> out_off;   // user job -> offset of its field
	std::vector<uint32_t> job_nv;              // user job -> vertices
	std::vector<uint32_t> slot_of;             // user job -> queue slot
	gp::BMesh *d_meshes = nullptr;
	gp::BJob *d_jobs = nullptr;
	uint32_t *d_queue = nullptr;
	float *d_out = nullptr;
	// Per-workgroup scratch, sized for the current configuration.
	uint32_t *d_level = nullptr, *d_sorted = nullptr, *d_limits = nullptr, *d_nlim = nullptr;
	unsigned long long *d_key = nullptr;
	float *d_dist = nullptr;
	uint32_t *d_fcid = nullptr, *d_fmidx = nullptr, *d_finv = nullptr;
	float2 *d_fgeo = nullptr;
	uint32_t njobs = 0, max_nv = 0, max_slots = 0, grid = 0, block = 0;
	unsigned long long total_out = 0;
	double tol = GP_PTP_TOL;
	hipStream_t stream = nullptr;
	hipGraphExec_t graph = nullptr;     // null when graphs are disabled
	gp::BArgs args{};                    // the persistent launch (without a graph)
	const void *fn = nullptr;
	hipEvent_t e0 = nullptr, e1 = nullptr;

	void free_scratch()
	{
		for (void *p : { (void *) d_level, (void *) d_sorted, (void *) d_limits, (void *) d_nlim, (void *) d_key,
		                 (void *) d_dist, (void *) d_fcid, (void *) d_fmidx, (void *) d_finv, (void *) d_fgeo })
			(void) hipFree(p);
		d_level = d_sorted = d_limits = d_nlim = nullptr;
		d_key = nullptr;
		d_dist = nullptr;
		d_fcid = d_fmidx = d_finv = nullptr;
		d_fgeo = nullptr;
		if (graph) (void) hipGraphExecDestroy(graph);
		graph = nullptr;
	}
	~gp_batch()
	{
		free_scratch();
		for (auto &t : tabs)
			for (void *p : { (void *) t.coff, (void *) t.cid, (void *) t.cgeo, (void *) t.cgeo64, (void *) t.perm })
				(void) hipFree(p);
		for (void *p : { (void *) d_meshes, (void *) d_jobs, (void *) d_queue, (void *) d_out }) (void) hipFree(p);
		if (e0) (void) hipEventDestroy(e0);
		if (e1) (void) hipEventDestroy(e1);
		if (stream) (void) hipStreamDestroy(stream);
	}
};

namespace gp {

template <typename V>
static int up(V *&dst, const V *src, size_t n)
{
	if (hipMalloc(&dst, sizeof(V) * std::max<size_t>(n, 1)) != hipSuccess) return fail(GP_ERR_NOMEM, "batch: allocation failed");
	if (n) GP_HIP(hipMemcpy(dst, src, sizeof(V) * n, hipMemcpyHostToDevice));
	return GP_OK;
}

// Fields whose fans all fit 8 slots run the 8-slot register path.
template <uint32_t NT>
static int launch_shape(uint32_t njobs, bool ms8, uint32_t &grid, const void *&fn)
{
	int dev = 0, ncu = 0, occ = 0;
	GP_HIP(hipGetDevice(&dev));
	GP_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
	fn = ms8 ? (const void *) k_batch<NT, 8> : (const void *) k_batch<NT, MAXS>;
	GP_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, fn, NT, 0));
	grid = std::max<uint32_t>(1, std::min<uint32_t>(njobs, (uint32_t) ncu * (uint32_t) std::max(1, occ)));
	return GP_OK;
}

// Sizes the per-workgroup scratch for `block` threads per field and captures
// the queue reset + persistent launch into the batch's graph.
static int batch_configure(gp_batch *b, uint32_t block)
{
	b->free_scratch();
	uint32_t grid = 0;
	const void *fn = nullptr;
	int rc;
	switch (block) {
	case 256: rc = launch_shape<256>(b->njobs, b->max_slots <= 8, grid, fn); break;
	case 512: rc = launch_shape<512>(b->njobs, b->max_slots <= 8, grid, fn); break;
	case 1024: rc = launch_shape<1024>(b->njobs, b->max_slots <= 8, grid, fn); break;
	default: return fail(GP_ERR_INVALID, "batch: block must be 256, 512 or 1024 (got %u)", block);
	}
	if (rc) return rc;
	const size_t mv = b->max_nv;
	// Field tables: the largest field's chunks at the longest fan's width.
	const size_t fslots = ((mv + 63) / 64) * 64 * (size_t) b->max_slots;
	const size_t per_wg = 4 * mv * 3 + 4 * (mv + 2) + 8 * mv + 8 * mv + fslots * (4 + 4 + sizeof(float2));
	// At most 32 GiB of scratch: a wider grid than that would take no more
	// fields at once than the workgroups that fit.
	const size_t cap = std::max<size_t>(1, (32ull << 30) / std::max<size_t>(per_wg, 1));
	grid = (uint32_t) std::min<size_t>(grid, cap);
	const size_t g = grid;
	if (hipMalloc(&b->d_level, 4 * mv * g) != hipSuccess || hipMalloc(&b->d_sorted, 4 * mv * g) != hipSuccess ||
	    hipMalloc(&b->d_limits, 4 * (mv + 2) * g) != hipSuccess || hipMalloc(&b->d_nlim, 4 * g) != hipSuccess ||
	    hipMalloc(&b->d_key, 8 * mv * g) != hipSuccess || hipMalloc(&b->d_dist, 8 * mv * g) != hipSuccess ||
	    hipMalloc(&b->d_finv, 4 * mv * g) != hipSuccess || hipMalloc(&b->d_fcid, 4 * fslots * g) != hipSuccess ||
	    hipMalloc(&b->d_fmidx, 4 * fslots * g) != hipSuccess || hipMalloc(&b->d_fgeo, sizeof(float2) * fslots * g) != hipSuccess) {
		b->free_scratch();
		return fail(GP_ERR_NOMEM, "batch: scratch allocation failed (%u workgroups x %u vertices)", grid, b->max_nv);
	}
	BArgs a{ b->d_meshes, b->d_jobs, b->njobs, b->max_nv, b->d_queue, b->d_level, b->d_sorted, b->d_limits, b->d_nlim,
	         b->d_key, b->d_dist, b->d_out, b->tol, b->d_fcid, b->d_fmidx, b->d_finv, b->d_fgeo, fslots };
	b->args = a;
	b->fn = fn;
	b->grid = grid;
	b->block = block;
	if (!graphs_enabled()) return GP_OK;
	void *kargs[] = { &a };
	uint32_t *queue = b->d_queue;
	void *rargs[] = { &queue };
	GraphChain gc;
	gc.add((const void *) k_batch_reset, dim3(1), dim3(1), rargs);
	gc.add(fn, dim3(grid), dim3(block), kargs);
	const hipError_t e = gc.instantiate(&b->graph);
	if (e != hipSuccess) return fail(GP_ERR_HIP, "batch graph: %s", hipGetErrorString(e));
	b->grid = grid;
	b->block = block;
	return GP_OK;
}

static int batch_launch(gp_batch *b, float *ms)
{
	GP_HIP(hipEventRecord(b->e0, b->stream));
	if (b->graph) {
		GP_HIP(hipGraphLaunch(b->graph, b->stream));
	} else {
		void *kargs[] = { &b->args };
		k_batch_reset<<<1, 1, 0, b->stream>>>(b->d_queue);
		GP_HIP(hipLaunchKernel(b->fn, dim3(b->grid), dim3(b->block), kargs, 0, b->stream));
	}
	GP_HIP(hipEventRecord(b->e1, b->stream));
	GP_HIP(hipStreamSynchronize(b->stream));
	GP_HIP(hipEventElapsedTime(ms, b->e0, b->e1));
	return GP_OK;
}

int batch_create(Che *const *meshes, uint32_t nmeshes, const uint32_t *job_mesh, const uint32_t *job_source,
                 uint32_t njobs, double tol, uint32_t block, gp_batch **out)
{
	*out = nullptr;
	if (!meshes || !nmeshes || !job_mesh || !job_source || !njobs) return fail(GP_ERR_INVALID, "batch: empty input");
	if (!(tol > 0)) return fail(GP_ERR_INVALID, "batch: tol must be positive");
	for (uint32_t k = 0; k < nmeshes; k++)
		if (!meshes[k]) return fail(GP_ERR_INVALID, "batch: mesh %u is null", k);
	for (uint32_t q = 0; q < njobs; q++) {
		if (job_mesh[q] >= nmeshes) return fail(GP_ERR_INVALID, "batch: job %u mesh out of range", q);
		if (job_source[q] >= meshes[job_mesh[q]]->nv) return fail(GP_ERR_INVALID, "batch: job %u source out of range", q);
	}
	auto b = std::make_unique<gp_batch>();
	b->njobs = njobs;
	b->tol = tol;
	std::vector<BMesh> bm(nmeshes);
	b->tabs.resize(nmeshes);
	for (uint32_t k = 0; k < nmeshes; k++) {
		const Che &m0 = *meshes[k];
		if (m0.nv >= SLOT_WRAP) return fail(GP_ERR_CAPACITY, "batch: mesh %u has too many vertices", k);
		for (uint32_t v = 0; v < m0.nv; v++)
			if (m0.fan_off[v + 1] - m0.fan_off[v] > MAX_LINK)
				return fail(GP_ERR_CAPACITY, "batch: mesh %u vertex %u: link over %u", k, v, MAX_LINK);
		auto &t = b->tabs[k];
		std::vector<uint32_t> order;
		morton_order(m0, order);
		t.ren = std::make_unique<Che>();
		int rc = coalesce(m0, order.data(), m0.nv, *t.ren, t.inv);
		if (rc) return rc;
		Che &m = *t.ren;
		m.mean_edge = m0.mean_edge;      // the same h as the mesh itself
		if ((rc = upload_mesh(m))) return rc;
		Layout L;
		if ((rc = build_layout(m, nullptr, m.nv, false, L, 0, 1, 0, true))) return rc;
		if ((rc = up(t.coff, L.coff.data(), L.coff.size())) || (rc = up(t.cid, L.cid.data(), L.cid.size())) ||
		    (rc = up(t.cgeo, reinterpret_cast<const float2 *>(L.cgeo.data()), L.cgeo.size())) ||
		    (rc = up(t.cgeo64, reinterpret_cast<const double4 *>(L.cgeo64.data()), L.cgeo64.size())) ||
		    (rc = up(t.perm, order.data(), order.size())))
			return rc;
		bm[k] = BMesh{ m.dev->fan_off, m.dev->fan, m.dev->closed, t.coff, t.cid, t.cgeo, t.cgeo64, nullptr, t.perm,
		               m.dev->h, m.dev->inv_h, m.nv, L.max_slots };
		b->max_nv = std::max(b->max_nv, m.nv);
		b->max_slots = std::max(b->max_slots, L.max_slots);
	}
	// Queue order: longest fields first (work of a field grows ~ V^1.5), so
	// the last jobs taken are the short ones and the grid drains evenly.
	std::vector<uint32_t> order(njobs);
	std::iota(order.begin(), order.end(), 0u);
	std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) {
		return meshes[job_mesh[x]]->nv > meshes[job_mesh[y]]->nv;
	});
	b->slot_of.resize(njobs);
	b->out_off.resize(njobs);
	b->job_nv.resize(njobs);
	std::vector<BJob> jobs(njobs);
	unsigned long long off = 0;
	for (uint32_t s = 0; s < njobs; s++) {
		const uint32_t q = order[s];
		const uint32_t nv = meshes[job_mesh[q]]->nv;
		b->slot_of[q] = s;
		b->out_off[q] = off;
		b->job_nv[q] = nv;
		jobs[s] = BJob{ job_mesh[q], b->tabs[job_mesh[q]].inv[job_source[q]], 0, 0, off, 0 };
		off += nv;
	}
	b->total_out = off;
	int rc;
	if ((rc = up(b->d_meshes, bm.data(), bm.size())) || (rc = up(b->d_jobs, jobs.data(), jobs.size()))) return rc;
	if (hipMalloc(&b->d_queue, 4) != hipSuccess || hipMalloc(&b->d_out, 4 * std::max<size_t>(off, 1)) != hipSuccess)
		return fail(GP_ERR_NOMEM, "batch: output allocation failed (%llu values)", off);
	GP_HIP(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
	GP_HIP(hipEventCreate(&b->e0));
	GP_HIP(hipEventCreate(&b->e1));
	if (block) {
		if ((rc = batch_configure(b.get(), block))) return rc;
	} else {
		// Autotune threads per field: one timed run of the whole batch per
		// candidate after one warm-up run, keep the fastest.
		float best = INFINITY;
		uint32_t best_block = 1024;
		for (uint32_t cand : { 256u, 512u, 1024u }) {
			float ms = 0;
			if ((rc = batch_configure(b.get(), cand)) || (rc = batch_launch(b.get(), &ms)) ||
			    (rc = batch_launch(b.get(), &ms)))
				return rc;
			if (ms < best) { best = ms; best_block = cand; }
		}
		if (b->block != best_block && (rc = batch_configure(b.get(), best_block))) return rc;
	}
	*out = b.release();
	return GP_OK;
}

int batch_run(gp_batch *b, gp_batch_stats *st)
{
	if (!b) return fail(GP_ERR_INVALID, "batch is null");
	float ms = 0;
	int rc = batch_launch(b, &ms);
	if (rc) return rc;
	if (st) {
		std::vector<BJob> jobs(b->njobs);
		GP_HIP(hipMemcpy(jobs.data(), b->d_jobs, sizeof(BJob) * b->njobs, hipMemcpyDeviceToHost));
		unsigned long long s = 0, r = 0;
		for (auto &j : jobs) { s += j.sweeps; r += j.relax; }
		st->sweeps = s;
		st->relaxations = r;
		st->vertex_fields = b->total_out;
		st->block = b->block;
		st->grid = b->grid;
		st->kernel_ms = ms;
	}
	return GP_OK;
}

int batch_result(const gp_batch *b, uint32_t job, double *dist, uint32_t cap, uint32_t *sweeps)
{
	if (!b || !dist) return fail(GP_ERR_INVALID, "null argument");
	if (job >= b->njobs) return fail(GP_ERR_INVALID, "batch: job %u out of range", job);
	const uint32_t n = b->job_nv[job];
	if (cap < n) return fail(GP_ERR_CAPACITY, "batch: result of job %u needs %u entries", job, n);
	std::vector<float> tmp(n);
	GP_HIP(hipMemcpy(tmp.data(), b->d_out + b->out_off[job], 4ull * n, hipMemcpyDeviceToHost));
	for (uint32_t v = 0; v < n; v++) dist[v] = tmp[v];
	if (sweeps) {
		BJob j;
		GP_HIP(hipMemcpy(&j, b->d_jobs + b->slot_of[job], sizeof(BJob), hipMemcpyDeviceToHost));
		*sweeps = j.sweeps;
	}
	return GP_OK;
}

}  // namespace gp

extern "C" void gp_batch_destroy(gp_batch *b) { delete b; }
