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
// pairs are built once into a hipGraph and replayed until the device sets
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
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <tuple>
#include <type_traits>
#include <vector>

#include "device.hpp"
#include "levels.hpp"
#include "sweep_dev.hpp"
#include "flow_kernel.hpp"
#include "sweep_kernel.hpp"
#include "update_step.hpp"

namespace gp {

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

// A mesh's device state (upload, vertex-major tables) is built on first use
// and kept; plans of one mesh may be created from several host threads at
// once, so that first use is serialised.
static std::recursive_mutex g_mesh_mu;

// Uploads the fan tables (ring builder) and FP64 positions once per mesh.
int upload_mesh(Che &m)
{
	std::lock_guard<std::recursive_mutex> lock(g_mesh_mu);
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
	bool band = false;                       // band schedule: final check sweeps over every position
	unsigned long long n_flagged = 0;
	gp::PatchPlan *pp = nullptr;             // patch / band schedules
	uint32_t *d_order = nullptr, *d_src = nullptr, *d_trace = nullptr;
	uint32_t *d_limits = nullptr, *d_coff = nullptr, *d_cid = nullptr;
	unsigned long long *d_fcum = nullptr;
	float2 *d_cgeo = nullptr;
	double4 *d_cgeo64 = nullptr;
	gp::Edge64 *d_cx64 = nullptr;
	void *d_dist0 = nullptr, *d_dist1 = nullptr;
	uint32_t *d_cl0 = nullptr, *d_cl1 = nullptr;   // cluster labels (opts.clusters)
	uint32_t *d_inv = nullptr;                // device-built plans: original vertex -> position (NIL: unreached)
	double *d_out = nullptr;
	double *d_orig = nullptr;                 // device-built plans: the result in the original numbering
	gp::PState *d_state = nullptr, *h_state = nullptr;   // h_state: 2 pinned snapshots
	hipStream_t stream = nullptr;
	hipEvent_t e0 = nullptr, e1 = nullptr, ev[2] = { nullptr, nullptr };
	hipGraphExec_t g_big = nullptr, g_small = nullptr;
	uint32_t big = 64, small = 8;            // sweeps per graph replay (multiples of 4)
	uint32_t xblk = 0;                       // log2 of the chunk blocks dealt to XCDs (0: eighths)
	bool has_run = false;
	bool capped = false;                     // the last run stopped at its sweep cap
	uint32_t check_sweeps = 0;               // check sweeps of the last run (band plans)
	// Dataflow launch (GP_LAUNCH_FLOW): eligibility (-1: not yet checked) and buffers.
	int flow_ok = -1;
	bool flow_last = false;                  // the last run was one dataflow launch
	uint32_t flow_grid = 0;
	unsigned long long *d_pr0 = nullptr, *d_pr1 = nullptr, *d_facc = nullptr;
	uint32_t *d_cflag = nullptr, *d_hist = nullptr;
	gp::FlowState *d_fs = nullptr, *h_fs = nullptr;

	~gp_plan()
	{
		for (void *q : { (void *) d_pr0, (void *) d_pr1, (void *) d_facc, (void *) d_cflag, (void *) d_hist,
		                 (void *) d_fs })
			(void) hipFree(q);
		if (h_fs) (void) hipHostFree(h_fs);
		for (void *p : { (void *) d_limits, (void *) d_order, (void *) d_src, (void *) d_trace, (void *) d_coff,
		                 (void *) d_cid, (void *) d_cgeo, (void *) d_cgeo64, (void *) d_cx64, d_dist0, d_dist1,
		                 (void *) d_out, (void *) d_state, (void *) d_fcum, (void *) d_cl0, (void *) d_cl1,
		                 (void *) d_inv, (void *) d_orig })
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

// The precision a plan runs in: FP32 requests on meshes with more than a
// quarter of their fan triangles flagged run FP64 sweeps (GP_GEO_AUTO).
static bool plan_f64(const Che &mesh, const gp_ptp_opts &opts)
{
	if (opts.precision == GP_PREC_F64) return true;
	return opts.fp32_geometry == GP_GEO_AUTO && prefer_f64(mesh);
}

static bool ring_like(int schedule)
{
	return schedule == GP_SCHED_RINGS || schedule == GP_SCHED_LEARNED_BAND || schedule == GP_SCHED_EUCLID_BAND ||
	       schedule == GP_SCHED_RING_BAND;
}

// Builds a plan from given rings (limits/sorted in the original numbering;
// for the band schedules, the bands in the order to sweep them).
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
	if (opts.precision != GP_PREC_F32 && opts.precision != GP_PREC_F64)
		return fail(GP_ERR_INVALID, "plan: unknown precision %d", opts.precision);
	if (!ring_like(opts.schedule)) return fail(GP_ERR_INVALID, "plan_create_rings: ring or band schedule only");
	auto p = std::make_unique<gp_plan>();
	p->mesh = mesh;
	p->opts = opts;
	p->opts.precision = plan_f64(*mesh, opts) ? GP_PREC_F64 : GP_PREC_F32;
	p->band = opts.schedule != GP_SCHED_RINGS && opts.band_check;
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
	const bool f64 = p->opts.precision == GP_PREC_F64;
	Layout L;
	rc = build_layout(*use, opts.coalesce ? nullptr : sorted, nr, f64, L, 0, 64, 0, true);
	if (rc) return rc;
	p->nchunks = L.nchunks;
	p->max_slots = L.max_slots;
	p->sb = L.uniform;
	p->n_flagged = L.n_flagged;
	p->g64 = !f64 && opts.fp32_geometry == GP_GEO_FP64;
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
		std::vector<float2> cg(L.cgeo.size());
		std::memcpy(cg.data(), L.cgeo.data(), sizeof(float2) * cg.size());
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
                                uint32_t nv, double ih2, uint32_t *cid, float2 *cgeo, double4 *cgeo64, Edge64 *cx64)
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
					// det = NaN where the reference's own determinant is 0
					// (its update is NaN there; see build_layout).
					const bool ref_degenerate = q0 * q1 - g * g == 0;
					if (ref_degenerate) gd.z = __longlong_as_double(0x7FF8000000000000ll);
					if (ref_degenerate || !(qq > 0) || det < FLAG_SIN2_MIN * qq || g < FLAG_COS_MIN * sqrt(qq) ||
					    q0 > FLAG_Q_RATIO * q1 || q1 > FLAG_Q_RATIO * q0)
						id |= F64_FLAG;
				}
				cgeo[at] = make_float2((float) gd.x, (float) gd.y);
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
	std::lock_guard<std::recursive_mutex> lock(g_mesh_mu);
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
		         : (hipMalloc(&m->cgeo, sizeof(float2) * nf) == hipSuccess ? hipMalloc(&m->cgeo64, sizeof(double4) * nf)
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
// Largest BFS level (NIL excluded) into *out (zeroed beforehand).
__global__ void k_max_level(const uint32_t *level, uint32_t n, uint32_t *out)
{
	uint32_t m = 0;
	for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x)
		if (level[v] != GP_NIL) m = max(m, level[v]);
	for (uint32_t o = 32; o > 0; o >>= 1) m = max(m, (uint32_t) __shfl_xor((int) m, (int) o));
	if ((threadIdx.x & 63u) == 0 && m) atomicMax(out, m);
}

int plan_create_device(Che *mesh, const uint32_t *src, uint32_t ns, const gp_ptp_opts &opts, gp_plan **out,
                       const double *dhat, double width, gp_mlp *mlp)
{
	*out = nullptr;
	if (!mesh || !src) return fail(GP_ERR_INVALID, "plan: null argument");
	if (ns == 0) return fail(GP_ERR_INVALID, "plan: no sources");
	if (!ring_like(opts.schedule) || !opts.coalesce || opts.clusters)
		return fail(GP_ERR_INVALID, "plan_create_device: ring or band schedule, coalesced, without labels only");
	if (opts.precision != GP_PREC_F32 && opts.precision != GP_PREC_F64)
		return fail(GP_ERR_INVALID, "plan: unknown precision %d", opts.precision);
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
	const bool f64 = plan_f64(*mesh, opts);
	DeviceMesh *dm = mesh->dev;
	VertexTables mt{};
	uint32_t S = 0;
	if ((rc = master_tables(*mesh, f64, mt, S))) return rc;
	if (S > MAXS) return fail(GP_ERR_CAPACITY, "plan: a fan needs %u slots", S);
	auto p = std::make_unique<gp_plan>();
	p->mesh = mesh;
	p->opts = opts;
	p->opts.precision = f64 ? GP_PREC_F64 : GP_PREC_F32;
	p->band = opts.schedule != GP_SCHED_RINGS && opts.band_check;
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
	if (dhat || mlp) {
		// Bands of dhat in place of the BFS levels of the reached vertices:
		// the caller's dhat, or the network's prediction from these levels
		// (the features' ring index), kept on the device.
		if (!(width > 0)) return fail(GP_ERR_INVALID, "band width must be positive");
		double *d_dhat = nullptr;
		uint32_t *d_bmax = nullptr;
		if (dhat) {
			std::vector<double> dh(dhat, dhat + mesh->nv);
			if ((rc = upload(d_dhat, dh))) return rc;
		} else if (hipMalloc(&d_dhat, sizeof(double) * mesh->nv) != hipSuccess) {
			return fail(GP_ERR_NOMEM, "plan: device allocation failed");
		}
		std::unique_ptr<double, hipError_t (*)(void *)> dhat_guard(d_dhat, hipFree);
		if (hipMalloc(&d_bmax, 4) != hipSuccess) return fail(GP_ERR_NOMEM, "plan: device allocation failed");
		std::unique_ptr<uint32_t, hipError_t (*)(void *)> bmax_guard(d_bmax, hipFree);
		const uint32_t gb = std::max<uint32_t>(1, std::min<uint32_t>(2048, (mesh->nv + 255) / 256));
		if (!dhat) {
			k_fill_u32<<<1, 64, 0, p->stream>>>(d_bmax, 1, 0);
			k_max_level<<<gb, 256, 0, p->stream>>>(lv.level, mesh->nv, d_bmax);
			uint32_t lmax = 0;
			GP_HIP(hipMemcpyAsync(&lmax, d_bmax, 4, hipMemcpyDeviceToHost, p->stream));
			GP_HIP(hipStreamSynchronize(p->stream));
			if ((rc = mlp_predict_device(mlp, *mesh, d_srcv, ns, lv.level, lmax, d_dhat, p->stream))) return rc;
		}
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
	    hipMalloc(&p->d_orig, sizeof(double) * mesh->nv) != hipSuccess ||
	    hipMalloc(&p->d_coff, 4ull * (p->nchunks + 1)) != hipSuccess ||
	    hipMalloc(&p->d_cid, 4 * std::max<size_t>(lanes, 1)) != hipSuccess)
		return fail(GP_ERR_NOMEM, "plan: device allocation failed");
	if (f64) {
		if (hipMalloc(&p->d_cx64, sizeof(Edge64) * std::max<size_t>(lanes, 1)) != hipSuccess)
			return fail(GP_ERR_NOMEM, "plan: device allocation failed");
	} else if (hipMalloc(&p->d_cgeo, sizeof(float2) * std::max<size_t>(lanes, 1)) != hipSuccess ||
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
	a.npos = p->n;
	a.xblk = p->xblk;
	a.cl0 = p->d_cl0; a.cl1 = p->d_cl1;
	return a;
}

// The plan's sweep kernel variant.
static const void *sweep_kernel(const gp_plan *p, bool cl)
{
	if (p->opts.precision == GP_PREC_F64) {
		if (cl) return p->ms8 ? (const void *) k_sweep<true, true, 8> : (const void *) k_sweep<true, true, MAXS>;
		return p->ms8 ? (const void *) k_sweep<true, false, 8> : (const void *) k_sweep<true, false, MAXS>;
	}
	if (p->g64) {
		if (cl) return p->ms8 ? (const void *) k_sweep<false, true, 8, true> : (const void *) k_sweep<false, true, MAXS, true>;
		return p->ms8 ? (const void *) k_sweep<false, false, 8, true> : (const void *) k_sweep<false, false, MAXS, true>;
	}
	if (cl) return p->ms8 ? (const void *) k_sweep<false, true, 8> : (const void *) k_sweep<false, true, MAXS>;
	return p->ms8 ? (const void *) k_sweep<false, false, 8> : (const void *) k_sweep<false, false, MAXS>;
}

// Launches the plan's sweep kernel variant with argument k4 (a sweep slot,
// or 4 + d for a check sweep from buffer d).
static void launch_sweep(const gp_plan *p, const KArgs &a, uint32_t k4)
{
	KArgs ka = a;
	PState *st = a.st;
	void *args[] = { &st, &k4, &ka };
	(void) hipLaunchKernel(sweep_kernel(p, a.cl0 != nullptr), dim3(p->grid), dim3(p->block), args, 0, p->stream);
}

// Enqueues sweeps k0 .. k0+count-1 (their k & 3 slots).
static void enqueue_sweeps(const gp_plan *p, const KArgs &a, uint32_t k0, uint32_t count)
{
	for (uint32_t q = 0; q < count; q++) launch_sweep(p, a, (k0 + q) & 3u);
}

// Sweep kernel variant bits of gp_ptp_stats.kernel.
static uint32_t kernel_bits(const gp_plan *p)
{
	return (p->opts.precision == GP_PREC_F64 ? 1u : 0u) | (p->d_cl0 ? 2u : 0u) | (p->ms8 ? 4u : 0u) |
	       (p->g64 && p->opts.precision != GP_PREC_F64 ? 8u : 0u) | (p->flow_last ? 16u : 0u);
}

static int capture(gp_plan *p, const KArgs &a, uint32_t count, hipGraphExec_t *out)
{
	KArgs ka = a;
	PState *st = a.st;
	const uint32_t k4s[4] = { 0, 1, 2, 3 };
	const void *fn = sweep_kernel(p, a.cl0 != nullptr);
	GraphChain gc;
	for (uint32_t q = 0; q < count; q++) {
		void *args[] = { &st, (void *) &k4s[q & 3u], &ka };
		gc.add(fn, dim3(p->grid), dim3(p->block), args);
	}
	const hipError_t e = gc.instantiate(out);
	if (e != hipSuccess) return fail(GP_ERR_HIP, "sweep graph: %s", hipGetErrorString(e));
	return GP_OK;
}

// The dataflow kernel variant of a plan.
static const void *flow_kernel(const gp_plan *p)
{
	if (p->g64) return p->ms8 ? (const void *) k_flow<8, true> : (const void *) k_flow<MAXS, true>;
	return p->ms8 ? (const void *) k_flow<8, false> : (const void *) k_flow<MAXS, false>;
}

// Whether the plan can run as one dataflow launch: FP32 without labels on the
// ring-numbered chunked layout, every fan in registers, every ring nonempty
// (a tested ring then has a chunk to publish the next state), at most
// FLOW_MAX_OWN chunks per resident wave.  Allocates the buffers on success.
static int flow_prepare(gp_plan *p, bool &ok)
{
	ok = false;
	if (p->flow_ok >= 0) {
		ok = p->flow_ok != 0;
		return GP_OK;
	}
	p->flow_ok = 0;
	if (p->opts.precision == GP_PREC_F64 || p->d_cl0 || p->d_order || p->pp || p->max_slots > MAXS || p->nl < 3)
		return GP_OK;
	std::vector<uint32_t> lim(p->nl);
	GP_HIP(hipMemcpy(lim.data(), p->d_limits, 4ull * p->nl, hipMemcpyDeviceToHost));
	for (uint32_t q = 1; q < p->nl; q++)
		if (lim[q] <= lim[q - 1]) return GP_OK;
	if (lim[p->nl - 1] != p->nr || p->nl >= FLOW_IMASK) return GP_OK;
	int dev = 0, ncu = 0, occ = 0;
	GP_HIP(hipGetDevice(&dev));
	GP_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
	GP_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, flow_kernel(p), 64, 0));
	const uint32_t waves = (uint32_t) ncu * (uint32_t) std::max(occ, 0);
	const uint32_t nch = (p->nr + 63) / 64;
	if (waves == 0 || nch > FLOW_MAX_OWN * (size_t) waves) return GP_OK;
	p->flow_grid = waves;
	const size_t nh = (size_t) p->max_iter + 2;
	if (hipMalloc(&p->d_pr0, 8ull * nch * 64) != hipSuccess || hipMalloc(&p->d_pr1, 8ull * nch * 64) != hipSuccess ||
	    hipMalloc(&p->d_facc, 8 * nh) != hipSuccess || hipMalloc(&p->d_cflag, 4ull * nch) != hipSuccess ||
	    hipMalloc(&p->d_hist, 4 * nh) != hipSuccess || hipMalloc(&p->d_fs, sizeof(FlowState)) != hipSuccess ||
	    hipHostMalloc(&p->h_fs, sizeof(FlowState)) != hipSuccess)
		return fail(GP_ERR_NOMEM, "plan: dataflow buffer allocation failed");
	p->flow_ok = 1;
	ok = true;
	return GP_OK;
}

// One run of the schedule as a single cooperative launch (after k_init and
// k_sources): pairs from the initial distances, the dataflow sweep, both
// distance buffers back from the pairs, and the state snapshot; the event
// ev[0] ends the timed region.
static int flow_run(gp_plan *p, const KArgs &a)
{
	const uint32_t nch = (p->nr + 63) / 64;
	const uint32_t nh = p->max_iter + 2;
	const uint32_t g = std::max<uint32_t>(1, std::min<uint32_t>(2048, (p->n + 255) / 256));
	k_flow_init<<<g, 256, 0, p->stream>>>((const float *) p->d_dist0, p->d_pr0, p->d_pr1, p->nr, p->d_cflag, nch,
	                                       p->d_hist, p->d_facc, nh, p->d_fs);
	GP_HIP(hipGetLastError());
	FlowArgs fa{};
	fa.a = a;
	fa.pr0 = p->d_pr0; fa.pr1 = p->d_pr1; fa.cflag = p->d_cflag; fa.hist = p->d_hist; fa.acc = p->d_facc;
	fa.fs = p->d_fs;
	fa.nchunks = nch;
	fa.nr = p->nr;
	// Wall-clock budget of the run (100 MHz ticks): 20 s, or
	// GPROSHAN_AMD_FLOW_DEADLINE_US microseconds (a short one makes the
	// bounded waits give up, which is how the failure path is tested).
	fa.budget = 2000000000ull;
	if (const char *e = std::getenv("GPROSHAN_AMD_FLOW_DEADLINE_US")) {
		const unsigned long long us = std::strtoull(e, nullptr, 10);
		if (us) fa.budget = std::min<unsigned long long>(us * 100ull, fa.budget);
	}
	void *args[] = { &fa };
	GP_HIP(hipLaunchCooperativeKernel(flow_kernel(p), dim3(p->flow_grid), dim3(64), args, 0, p->stream));
	k_flow_unpack<<<g, 256, 0, p->stream>>>(p->d_pr0, p->d_pr1, (float *) p->d_dist0, (float *) p->d_dist1, p->nr);
	snapshot(p->d_state, p->h_state, p->stream);
	snapshot(p->d_fs, p->h_fs, p->stream);
	GP_HIP(hipGetLastError());
	GP_HIP(hipEventRecord(p->ev[0], p->stream));
	GP_HIP(hipStreamSynchronize(p->stream));
	if (p->h_fs->abort) return fail(GP_ERR_KERNEL, "ptp: the dataflow sweep passed its deadline");
	if (!p->h_state->done) return fail(GP_ERR_KERNEL, "ptp: the dataflow sweep ended before the schedule");
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
	const uint32_t BIG = p->big, SMALL = p->small;
	bool flow = false;
	if (p->opts.launch_mode == GP_LAUNCH_FLOW) {
		int rc = flow_prepare(p, flow);
		if (rc) return rc;
	}
	p->flow_last = flow;
	const bool graph = !flow && p->opts.launch_mode == GP_LAUNCH_GRAPH && graphs_enabled();
	if (graph && !p->g_big) {
		int rc = capture(p, a, BIG, &p->g_big);
		if (!rc) rc = capture(p, a, SMALL, &p->g_small);
		if (rc) return rc;
	}
	p->has_run = false;
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
	// each replay a one-wave kernel writes the state into the replay's pinned
	// snapshot and an event is recorded; the host waits for the older
	// replay's event while the newer one runs, so the GPU does not idle on
	// the poll.  The device time ends at the event of the replay whose state
	// says done (a replay queued behind it only meets done slots and returns
	// at once).  One sweep more than the sweep count runs (it records the
	// last sweep's outcome).  j grows by one ring per sweep until |limits|-1,
	// so at least nl-1-j sweeps remain: big graphs while that bound allows,
	// small ones near the end.
	uint32_t launched = 0, j_seen = 2;
	const uint32_t cap = p->max_iter;
	auto replay = [&](uint32_t s) -> int {
		const bool big = p->nl - 1 > j_seen + BIG + (launched ? BIG : 0);
		const uint32_t cnt = big ? BIG : SMALL;
		if (graph) GP_HIP(hipGraphLaunch(big ? p->g_big : p->g_small, p->stream));
		else enqueue_sweeps(p, a, launched, cnt);
		launched += cnt;
		snapshot(p->d_state, p->h_state + s, p->stream);
		GP_HIP(hipGetLastError());
		GP_HIP(hipEventRecord(p->ev[s], p->stream));
		return GP_OK;
	};
	uint32_t cur = 0;
	bool done = done0 != 0;
	if (!done && flow) {
		int rc = flow_run(p, a);
		if (rc) return rc;
		done = true;
	} else if (!done) {
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
		snapshot(p->d_state, p->h_state + cur, p->stream);
		GP_HIP(hipGetLastError());
	}
	GP_HIP(hipEventRecord(p->e1, p->stream));
	GP_HIP(hipStreamSynchronize(p->stream));
	if (cur != 0) std::memcpy(p->h_state, p->h_state + cur, sizeof(PState));
	// The loop stopped at its cap when the last sweep's counter reached it
	// with ring i still below j.
	{
		const Slot &last = p->h_state->slot[3];
		p->capped = !done0 && last.iter >= p->max_iter && last.i < last.j;
	}
	// Band schedules: check sweeps over every position until one of them
	// moves no vertex by tol or more (a prediction that put a vertex in a
	// band ahead of its upwind neighbours cannot leave it unconverged).
	p->check_sweeps = 0;
	uint32_t fd = p->h_state->final_d;
	if (p->band) {
		const uint32_t check_cap = 4 * p->nl + 64;
		for (;;) {
			k_fill_u32<<<1, 64, 0, p->stream>>>(&p->d_state->check_bad, 1, 0);
			launch_sweep(p, a, 4u + fd);
			snapshot(p->d_state, p->h_state + 1, p->stream);
			GP_HIP(hipGetLastError());
			GP_HIP(hipStreamSynchronize(p->stream));
			fd ^= 1u;
			p->check_sweeps++;
			if (p->h_state[1].check_bad == 0) break;
			if (p->check_sweeps >= check_cap) {
				p->h_state->final_d = fd;
				return fail(GP_ERR_NOT_CONVERGED, "ptp: %u vertices still move after %u check sweeps",
				            p->h_state[1].check_bad, p->check_sweeps);
			}
		}
		GP_HIP(hipEventRecord(p->e1, p->stream));
		GP_HIP(hipStreamSynchronize(p->stream));
		p->h_state->final_d = fd;
	}
	float ms = 0;
	GP_HIP(hipEventElapsedTime(&ms, p->e0, p->band || done0 ? p->e1 : p->ev[cur]));
	p->has_run = true;
	if (stats) {
		stats->sweeps = p->h_state->sweeps;
		stats->n_limits = p->nl;
		stats->relaxations = p->h_state->relax + (unsigned long long) p->check_sweeps * p->n;
		stats->cond_checks = p->h_state->cond + (unsigned long long) p->check_sweeps * p->n;
		stats->f64_updates = p->h_state->f64;
		stats->patch_visits = 0;
		stats->deferred = 0;
		stats->kernel_ms = ms;
		stats->total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
		stats->precision = p->opts.precision;
		stats->capped = p->capped ? 1 : 0;
		stats->kernel = kernel_bits(p);
		stats->check_sweeps = p->check_sweeps;
	}
	// The uncapped loop has no cap of its own: reaching the guard means the
	// field did not converge (band plans were settled by the check sweeps).
	if (p->capped && p->opts.uncapped && !p->band)
		return fail(GP_ERR_NOT_CONVERGED, "ptp: the uncapped loop reached its guard of %u sweeps", p->max_iter);
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
		const uint32_t go = std::max<uint32_t>(1, std::min<uint32_t>(1024, (p->nv_orig + 255) / 256));
		k_from_positions<<<go, 256, 0, p->stream>>>(p->d_out, p->d_inv, p->nv_orig, p->d_orig);
		GP_HIP(hipGetLastError());
		GP_HIP(hipMemcpyAsync(dist, p->d_orig, sizeof(double) * p->nv_orig, hipMemcpyDeviceToHost, p->stream));
		GP_HIP(hipStreamSynchronize(p->stream));
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

// Autotuner of a plan's launch shape: the sweep kernel's workgroups (a
// quarter to all of those that fit on the device at once) and the sweeps per
// graph replay, timed on the plan itself (one warm-up run, then the
// mean device time of two) and remembered per plan shape, so that a later plan
// of the same shape takes the choice without timing.
namespace {
// The schedule is part of the key: learned and Euclidean bands of one mesh
// have different window sequences, and their best XCD deal differs.
struct TuneKey {
	uint32_t nv_log2, max_slots, f64, band, coalesce, uncapped, schedule;
	bool operator<(const TuneKey &o) const
	{
		return std::tie(nv_log2, max_slots, f64, band, coalesce, uncapped, schedule) <
		       std::tie(o.nv_log2, o.max_slots, o.f64, o.band, o.coalesce, o.uncapped, o.schedule);
	}
};
struct WidthKey {
	uint32_t nv_log2, max_slots, schedule, f64;
	bool operator<(const WidthKey &o) const
	{
		return std::tie(nv_log2, max_slots, schedule, f64) < std::tie(o.nv_log2, o.max_slots, o.schedule, o.f64);
	}
};
std::mutex g_tune_mu;
std::map<TuneKey, gp_tune> g_tune_cache;
std::map<WidthKey, double> g_width_cache;
bool g_tune_loaded = false;

// The choices persist across processes when GPROSHAN_AMD_TUNE_FILE names a
// file: it is read on first use and every new choice is appended, one line
// each ("shape <key> <grid> <replay> <xcd block> <ms>", "width <key> <width>");
// a line that does not parse whole (another key layout, a damaged entry) is
// skipped.
const char *tune_file() { return std::getenv("GPROSHAN_AMD_TUNE_FILE"); }

void tune_load_locked()
{
	if (g_tune_loaded) return;
	g_tune_loaded = true;
	const char *path = tune_file();
	if (!path || !*path) return;
	FILE *f = std::fopen(path, "r");
	if (!f) return;
	char line[512];
	while (std::fgets(line, sizeof(line), f)) {
		TuneKey k{};
		gp_tune t{};
		WidthKey wk{};
		double w = 0;
		int used = 0;
		const size_t len = std::strlen(line);
		if (std::sscanf(line, "shape %u %u %u %u %u %u %u %u %u %u %lf %n", &k.nv_log2, &k.max_slots, &k.f64,
		                &k.band, &k.coalesce, &k.uncapped, &k.schedule, &t.grid, &t.replay, &t.xcd_block,
		                &t.kernel_ms, &used) == 11 && (size_t) used == len) {
			// Only shapes the tuner itself can choose (a damaged file must not
			// make a plan capture an oversized graph).
			if (t.grid && t.grid <= (1u << 20) && t.replay >= 4 && t.replay <= 1024 && t.replay % 4 == 0 &&
			    t.xcd_block < 16)
				g_tune_cache[k] = t;
		} else if (std::sscanf(line, "width %u %u %u %u %lf %n", &wk.nv_log2, &wk.max_slots, &wk.schedule, &wk.f64,
		                       &w, &used) == 5 && (size_t) used == len) {
			if (w > 0 && w <= 64) g_width_cache[wk] = w;
		}
	}
	std::fclose(f);
}

void tune_append(const char *line)
{
	const char *path = tune_file();
	if (!path || !*path) return;
	if (FILE *f = std::fopen(path, "a")) {
		std::fputs(line, f);
		std::fclose(f);
	}
}

uint32_t log2_floor(uint32_t n)
{
	uint32_t lg = 0;
	while ((2ull << lg) <= n) lg++;
	return lg;
}

uint32_t mesh_max_slots(const Che &m)
{
	uint32_t s = 0;
	for (uint32_t v = 0; v < m.nv; v++) s = std::max(s, m.link_size(v));
	return s;
}
}  // namespace

bool tune_width_lookup(const Che &mesh, int schedule, int precision, double *width)
{
	const WidthKey k{ log2_floor(mesh.nv), mesh_max_slots(mesh), (uint32_t) schedule,
	                  precision == GP_PREC_F64 ? 1u : 0u };
	std::lock_guard<std::mutex> lock(g_tune_mu);
	tune_load_locked();
	auto it = g_width_cache.find(k);
	if (it == g_width_cache.end()) return false;
	*width = it->second;
	return true;
}

void tune_width_store(const Che &mesh, int schedule, int precision, double width)
{
	const WidthKey k{ log2_floor(mesh.nv), mesh_max_slots(mesh), (uint32_t) schedule,
	                  precision == GP_PREC_F64 ? 1u : 0u };
	std::lock_guard<std::mutex> lock(g_tune_mu);
	tune_load_locked();
	g_width_cache[k] = width;
	char line[160];
	std::snprintf(line, sizeof(line), "width %u %u %u %u %.17g\n", k.nv_log2, k.max_slots, k.schedule, k.f64, width);
	tune_append(line);
}

static void plan_drop_graphs(gp_plan *p)
{
	if (p->g_big) (void) hipGraphExecDestroy(p->g_big);
	if (p->g_small) (void) hipGraphExecDestroy(p->g_small);
	p->g_big = p->g_small = nullptr;
}

int plan_autotune(gp_plan *p, gp_tune *out)
{
	if (!p) return fail(GP_ERR_INVALID, "autotune: plan is null");
	if (p->pp) return fail(GP_ERR_INVALID, "autotune: ring and band plans only");
	const TuneKey key{ log2_floor(p->n), p->max_slots, p->opts.precision == GP_PREC_F64 ? 1u : 0u, p->band ? 1u : 0u,
	                   (uint32_t) (p->opts.coalesce != 0), (uint32_t) (p->opts.uncapped != 0),
	                   (uint32_t) p->opts.schedule };
	{
		std::lock_guard<std::mutex> lock(g_tune_mu);
		tune_load_locked();
		auto it = g_tune_cache.find(key);
		if (it != g_tune_cache.end()) {
			plan_drop_graphs(p);
			p->grid = it->second.grid;
			p->big = it->second.replay;
			p->xblk = it->second.xcd_block;
			if (out) { *out = it->second; out->candidates = 0; }
			return GP_OK;
		}
	}
	// The grid sweep_grid picked is the largest candidate (all resident
	// workgroups, or fewer when the plan has fewer chunks).
	const uint32_t gmax = p->grid;
	std::vector<uint32_t> grids;
	for (uint32_t q = 1; q <= 4; q++) grids.push_back(std::max<uint32_t>(8, (gmax * q / 4 + 7) & ~7u));
	grids.erase(std::unique(grids.begin(), grids.end()), grids.end());
	const uint32_t replays[] = { 16, 32, 64, 128 };
	gp_tune best{};
	best.kernel_ms = 1e300;
	uint32_t tried = 0;
	auto time_shape = [&](uint32_t g, uint32_t r, uint32_t xb) -> int {
		plan_drop_graphs(p);
		p->grid = g;
		p->big = r;
		p->xblk = xb;
		gp_ptp_stats st{};
		int rc = plan_run(p, &st);
		double ms = 0;
		for (int k = 0; k < 2 && !rc; k++) {
			rc = plan_run(p, &st);
			ms += st.kernel_ms / 2;
		}
		if (rc) return rc;
		tried++;
		if (ms < best.kernel_ms) { best.kernel_ms = ms; best.grid = g; best.replay = r; best.xcd_block = xb; }
		return GP_OK;
	};
	for (uint32_t g : grids)
		for (uint32_t r : replays)
			if (int rc = time_shape(g, r, 0)) return rc;
	// Then the XCD deal of the chosen shape: contiguous eighths of the window,
	// or blocks of 16, 64 or 256 chunks dealt round-robin to the XCDs.
	const uint32_t g_best = best.grid, r_best = best.replay;
	for (uint32_t xb : { 4u, 6u, 8u })
		if (int rc = time_shape(g_best, r_best, xb)) return rc;
	plan_drop_graphs(p);
	p->grid = best.grid;
	p->big = best.replay;
	p->xblk = best.xcd_block;
	best.candidates = tried;
	{
		std::lock_guard<std::mutex> lock(g_tune_mu);
		g_tune_cache[key] = best;
		char line[200];
		std::snprintf(line, sizeof(line), "shape %u %u %u %u %u %u %u %u %u %u %.6f\n", key.nv_log2, key.max_slots,
		              key.f64, key.band, key.coalesce, key.uncapped, key.schedule, best.grid, best.replay,
		              best.xcd_block, best.kernel_ms);
		tune_append(line);
	}
	if (out) *out = best;
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
		snapshot(p->d_state, p->h_state, p->stream);
		GP_HIP(hipGetLastError());
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
