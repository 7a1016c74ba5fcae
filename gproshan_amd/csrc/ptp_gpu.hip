// gfx950 parallel toplesets propagation: device mesh tables, the window
// sweep kernels and the device-resident solver plan.
//
// Schedule (src/geodesics_ptp.cpp:61-109; src/cuda/geodesics_ptp.cu:198-231):
// rings [i, j) form the window, i is raised to j/2 when it lags (K/2 band
// limit), ring i retires when every vertex in it changed by less than tol
// (relative), j grows by one ring per sweep, the double buffer flips each
// sweep, and at most 2*|limits| sweeps run.  Unlike the reference, which
// launches a relax kernel, a relative-error kernel and a Thrust count per
// sweep with three host synchronisations, the schedule state lives in device
// memory: each sweep kernel reads it, relaxes its window, counts
// non-converged ring-i vertices with workgroup reductions and one atomic per
// workgroup, and the last workgroup to finish advances the schedule.  The
// host only polls a "done" flag every few dozen sweeps.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <type_traits>

#include "device.hpp"
#include "update_step.hpp"

namespace gp {

void device_mesh_free(DeviceMesh *d)
{
	if (!d) return;
	(void) hipFree(d->fan_off);
	(void) hipFree(d->fan);
	(void) hipFree(d->geo);
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

// Host precompute of the FP32 triangle geometry and the FP64 flags, then upload.
int upload_mesh(Che &m)
{
	if (m.dev) return GP_OK;
	int rc = device_check(nullptr, 0);
	if (rc) return rc;
	const uint32_t nv = m.nv, nfan = m.fan_off[nv];
	std::vector<uint32_t> fan;
	std::vector<Geo32> geo;
	uint32_t flagged = 0;
	build_geo32(m, fan, geo, flagged);
	const double h = m.mean_edge, ih = 1.0 / h;
	auto dm = std::make_unique<DeviceMesh>();
	dm->nv = nv; dm->nfan = nfan; dm->n_flagged = flagged;
	dm->h = (float) h; dm->inv_h = (float) ih;
	auto cleanup = [&](int r) { device_mesh_free(dm.release()); return r; };
	if (hipMalloc(&dm->fan_off, sizeof(uint32_t) * (nv + 1)) != hipSuccess ||
	    hipMalloc(&dm->fan, sizeof(uint32_t) * std::max<uint32_t>(nfan, 1)) != hipSuccess ||
	    hipMalloc(&dm->geo, sizeof(float4) * std::max<uint32_t>(nfan, 1)) != hipSuccess ||
	    hipMalloc(&dm->pos, sizeof(double) * 3 * (size_t) nv) != hipSuccess ||
	    hipMalloc(&dm->closed, nv) != hipSuccess)
		return cleanup(fail(GP_ERR_NOMEM, "device mesh allocation failed"));
	if (hipMemcpy(dm->fan_off, m.fan_off.data(), sizeof(uint32_t) * (nv + 1), hipMemcpyHostToDevice) != hipSuccess ||
	    (nfan && hipMemcpy(dm->fan, fan.data(), sizeof(uint32_t) * nfan, hipMemcpyHostToDevice) != hipSuccess) ||
	    (nfan && hipMemcpy(dm->geo, geo.data(), sizeof(Geo32) * nfan, hipMemcpyHostToDevice) != hipSuccess) ||
	    hipMemcpy(dm->pos, m.gt.data(), sizeof(double) * 3 * (size_t) nv, hipMemcpyHostToDevice) != hipSuccess ||
	    hipMemcpy(dm->closed, m.fan_closed.data(), nv, hipMemcpyHostToDevice) != hipSuccess)
		return cleanup(fail(GP_ERR_HIP, "device mesh upload failed"));
	m.dev = dm.release();
	return GP_OK;
}

// ---------------------------------------------------------------------------
// Kernels

struct State {
	uint32_t i, j, d, iter, done, bad, arrived, sweeps;
	unsigned long long relax, cond, f64;
};

struct SweepArgs {
	State *st;
	const uint32_t *limits;
	const uint32_t *order;     // ring order; null when the mesh is numbered by rings
	const uint32_t *fan_off;
	const uint32_t *fan;
	const float4 *geo;
	const double *pos;
	void *dist0, *dist1;
	uint32_t *trace;           // 8 words per sweep, or null
	double tol;
	float h, inv_h;
	uint32_t nl, max_iter;
};

__device__ __forceinline__ float relax_f32(const SweepArgs &a, const float *od, uint32_t v, float best,
                                           uint32_t &f64n)
{
	const uint32_t s = a.fan_off[v], e = a.fan_off[v + 1];
	if (e - s < 2) return best;
	uint32_t ida = a.fan[s];
	uint32_t va = ida & ID_MASK;
	float ta = od[va];
	float4 ga = a.geo[s];
	for (uint32_t k = s + 1; k < e; k++) {
		const uint32_t idb = a.fan[k];
		const uint32_t vb = idb & ID_MASK;
		const float tb = od[vb];
		const float4 gb = a.geo[k];
		float p;
		if (ida & F64_FLAG) {
			p = (float) update_ref_f64(a.pos + 3 * (size_t) v, a.pos + 3 * (size_t) va, a.pos + 3 * (size_t) vb,
			                           (double) ta, (double) tb);
			f64n++;
		} else {
			p = update_stable_f32(ga.x, gb.x, ga.y, ga.z, ga.w, ta, tb, a.h, a.inv_h);
		}
		if (p < best) best = p;
		ida = idb; va = vb; ta = tb; ga = gb;
	}
	return best;
}

__device__ __forceinline__ double relax_f64(const SweepArgs &a, const double *od, uint32_t v, double best)
{
	const uint32_t s = a.fan_off[v], e = a.fan_off[v + 1];
	if (e - s < 2) return best;
	const double *gv = a.pos + 3 * (size_t) v;
	uint32_t va = a.fan[s] & ID_MASK;
	double ta = od[va];
	for (uint32_t k = s + 1; k < e; k++) {
		const uint32_t vb = a.fan[k] & ID_MASK;
		const double tb = od[vb];
		const double p = update_ref_f64(gv, a.pos + 3 * (size_t) va, a.pos + 3 * (size_t) vb, ta, tb);
		if (p < best) best = p;
		va = vb; ta = tb;
	}
	return best;
}

template <bool F64>
__global__ __launch_bounds__(256) void k_sweep(SweepArgs a)
{
	using T = typename std::conditional<F64, double, float>::type;
	State *st = a.st;
	if (st->done) return;
	uint32_t i = st->i;
	const uint32_t j = st->j, d = st->d;
	if (i < (j >> 1)) i = j >> 1;                   // K/2 limit band size
	const uint32_t start = a.limits[i], end = a.limits[j], cend = a.limits[i + 1];
	const T *od = (const T *) (d ? a.dist1 : a.dist0);
	T *nd = (T *) (d ? a.dist0 : a.dist1);

	uint32_t bad = 0, f64n = 0;
	const uint32_t stride = gridDim.x * blockDim.x;
	for (uint32_t k = start + blockIdx.x * blockDim.x + threadIdx.x; k < end; k += stride) {
		const uint32_t v = a.order ? a.order[k] : k;
		const T old = od[v];
		T best;
		if constexpr (F64) best = relax_f64(a, od, v, old);
		else best = relax_f32(a, od, v, old, f64n);
		nd[v] = best;
		if (k < cend) bad += !(fabs((double) best - (double) old) / (double) old < a.tol);
	}

	__shared__ uint32_t s_bad, s_f64;
	if (threadIdx.x == 0) { s_bad = 0; s_f64 = 0; }
	__syncthreads();
	if (bad) atomicAdd(&s_bad, bad);
	if (f64n) atomicAdd(&s_f64, f64n);
	__syncthreads();
	if (threadIdx.x != 0) return;
	if (s_bad) atomicAdd(&st->bad, s_bad);
	if (s_f64) atomicAdd(&st->f64, (unsigned long long) s_f64);
	__threadfence();
	const uint32_t prev = atomicAdd(&st->arrived, 1u);
	if (prev != gridDim.x - 1) return;
	// Last workgroup of this sweep: advance the schedule.
	__threadfence();
	const uint32_t nbad = atomicAdd(&st->bad, 0u);
	const uint32_t ncond = cend - start;
	if (a.trace) {
		uint32_t *r = a.trace + 8 * (size_t) st->sweeps;
		r[0] = st->sweeps + 1; r[1] = i; r[2] = j; r[3] = start;
		r[4] = end; r[5] = ncond; r[6] = ncond - nbad; r[7] = 0;
	}
	st->sweeps += 1;
	st->relax += end - start;
	st->cond += ncond;
	const uint32_t ni = i + (nbad == 0 ? 1u : 0u);
	const uint32_t nj = j + (j < a.nl - 1 ? 1u : 0u);
	const uint32_t it = st->iter + 1;
	st->i = ni; st->j = nj; st->d = d ^ 1u; st->iter = it;
	st->done = !(ni < nj) || !(it < a.max_iter);
	st->bad = 0;
	st->arrived = 0;
}

template <typename T>
__global__ void k_init(T *d0, T *d1, uint32_t n, const uint32_t *src, uint32_t ns, State *st, uint32_t done)
{
	const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
	for (uint32_t v = t; v < n; v += gridDim.x * blockDim.x) { d0[v] = (T) INFINITY; d1[v] = (T) INFINITY; }
	if (t == 0) {
		st->i = 1; st->j = 2; st->d = 0; st->iter = 0; st->done = done;
		st->bad = 0; st->arrived = 0; st->sweeps = 0; st->relax = 0; st->cond = 0; st->f64 = 0;
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
	uint32_t n = 0, nv_orig = 0, nl = 0, ns = 0, max_iter = 0, grid = 0, block = 256;
	uint32_t *d_limits = nullptr, *d_order = nullptr, *d_src = nullptr, *d_trace = nullptr;
	void *d_dist0 = nullptr, *d_dist1 = nullptr;
	double *d_out = nullptr;
	gp::State *d_state = nullptr, *h_state = nullptr;
	hipStream_t stream = nullptr;
	hipEvent_t e0 = nullptr, e1 = nullptr;
	bool has_run = false;

	~gp_plan()
	{
		(void) hipFree(d_limits); (void) hipFree(d_order); (void) hipFree(d_src); (void) hipFree(d_trace);
		(void) hipFree(d_dist0); (void) hipFree(d_dist1); (void) hipFree(d_out); (void) hipFree(d_state);
		if (h_state) (void) hipHostFree(h_state);
		if (e0) (void) hipEventDestroy(e0);
		if (e1) (void) hipEventDestroy(e1);
		if (stream) (void) hipStreamDestroy(stream);
	}
};

namespace gp {

// Builds a plan from given rings (limits/sorted in the original numbering).
int plan_create_rings(Che *mesh, const uint32_t *src, uint32_t ns, const uint32_t *limits, uint32_t nl,
                      const uint32_t *sorted, const gp_ptp_opts &opts, gp_plan **out)
{
	*out = nullptr;
	if (!mesh || !src || !limits || !sorted) return fail(GP_ERR_INVALID, "plan: null argument");
	if (ns == 0 || nl < 2) return fail(GP_ERR_INVALID, "plan: no sources / empty rings");
	auto p = std::make_unique<gp_plan>();
	p->mesh = mesh;
	p->opts = opts;
	p->nv_orig = mesh->nv;
	p->nl = nl;
	p->ns = ns;
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
	std::vector<uint32_t> srcm(src, src + ns);
	Che *use = mesh;
	if (opts.coalesce) {
		p->coal = std::make_unique<Che>();
		int rc = coalesce(*mesh, sorted, nr, *p->coal, p->inv);
		if (rc) return rc;
		for (auto &s : srcm) {
			if (s >= mesh->nv || p->inv[s] == NIL) return fail(GP_ERR_INVALID, "plan: source not in ring 0");
			s = p->inv[s];
		}
		use = p->coal.get();
	}
	int rc = upload_mesh(*use);
	if (rc) return rc;
	p->dm = use->dev;
	p->n = use->nv;
	p->max_iter = opts.max_sweeps ? opts.max_sweeps : (nl << 1);
	p->block = 256;
	p->grid = opts.grid ? opts.grid : std::max<uint32_t>(1, std::min<uint32_t>(1024, (p->n + 255) / 256));
	const bool f64 = opts.precision == GP_PREC_F64;
	const size_t es = f64 ? sizeof(double) : sizeof(float);
	GP_HIP(hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking));
	GP_HIP(hipEventCreate(&p->e0));
	GP_HIP(hipEventCreate(&p->e1));
	if (hipMalloc(&p->d_limits, sizeof(uint32_t) * nl) != hipSuccess ||
	    hipMalloc(&p->d_src, sizeof(uint32_t) * ns) != hipSuccess ||
	    hipMalloc(&p->d_dist0, es * p->n) != hipSuccess ||
	    hipMalloc(&p->d_dist1, es * p->n) != hipSuccess ||
	    hipMalloc(&p->d_out, sizeof(double) * p->n) != hipSuccess ||
	    hipMalloc(&p->d_state, sizeof(State)) != hipSuccess ||
	    hipMalloc(&p->d_trace, sizeof(uint32_t) * 8 * (size_t) p->max_iter) != hipSuccess ||
	    hipHostMalloc(&p->h_state, sizeof(State)) != hipSuccess)
		return fail(GP_ERR_NOMEM, "plan: device allocation failed");
	GP_HIP(hipMemcpy(p->d_limits, limits, sizeof(uint32_t) * nl, hipMemcpyHostToDevice));
	GP_HIP(hipMemcpy(p->d_src, srcm.data(), sizeof(uint32_t) * ns, hipMemcpyHostToDevice));
	if (!opts.coalesce) {
		GP_HIP(hipMalloc(&p->d_order, sizeof(uint32_t) * std::max<uint32_t>(nr, 1)));
		GP_HIP(hipMemcpy(p->d_order, sorted, sizeof(uint32_t) * nr, hipMemcpyHostToDevice));
	}
	*out = p.release();
	return GP_OK;
}

int plan_run(gp_plan *p, gp_ptp_stats *stats)
{
	if (!p) return fail(GP_ERR_INVALID, "plan is null");
	const auto t0 = std::chrono::steady_clock::now();
	const bool f64 = p->opts.precision == GP_PREC_F64;
	const uint32_t done0 = p->nl < 3 ? 1u : 0u;
	const uint32_t ig = std::min<uint32_t>(1024, (p->n + 255) / 256);
	if (f64) {
		k_init<double><<<ig, 256, 0, p->stream>>>((double *) p->d_dist0, (double *) p->d_dist1, p->n, p->d_src, p->ns, p->d_state, done0);
		k_sources<double><<<(p->ns + 255) / 256, 256, 0, p->stream>>>((double *) p->d_dist0, (double *) p->d_dist1, p->d_src, p->ns);
	} else {
		k_init<float><<<ig, 256, 0, p->stream>>>((float *) p->d_dist0, (float *) p->d_dist1, p->n, p->d_src, p->ns, p->d_state, done0);
		k_sources<float><<<(p->ns + 255) / 256, 256, 0, p->stream>>>((float *) p->d_dist0, (float *) p->d_dist1, p->d_src, p->ns);
	}
	GP_HIP(hipGetLastError());
	SweepArgs a{};
	a.st = p->d_state; a.limits = p->d_limits; a.order = p->d_order;
	a.fan_off = p->dm->fan_off; a.fan = p->dm->fan; a.geo = p->dm->geo; a.pos = p->dm->pos;
	a.dist0 = p->d_dist0; a.dist1 = p->d_dist1; a.trace = p->d_trace;
	a.tol = p->opts.tol; a.h = p->dm->h; a.inv_h = p->dm->inv_h; a.nl = p->nl; a.max_iter = p->max_iter;

	GP_HIP(hipEventRecord(p->e0, p->stream));
	const uint32_t chunk = 64;
	for (uint32_t launched = 0; launched < p->max_iter && !done0;) {
		const uint32_t c = std::min(chunk, p->max_iter - launched);
		for (uint32_t q = 0; q < c; q++) {
			if (f64) k_sweep<true><<<p->grid, p->block, 0, p->stream>>>(a);
			else k_sweep<false><<<p->grid, p->block, 0, p->stream>>>(a);
		}
		launched += c;
		GP_HIP(hipGetLastError());
		GP_HIP(hipMemcpyAsync(p->h_state, p->d_state, sizeof(State), hipMemcpyDeviceToHost, p->stream));
		GP_HIP(hipStreamSynchronize(p->stream));
		if (p->h_state->done) break;
	}
	GP_HIP(hipEventRecord(p->e1, p->stream));
	GP_HIP(hipMemcpyAsync(p->h_state, p->d_state, sizeof(State), hipMemcpyDeviceToHost, p->stream));
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
	const uint32_t d = p->h_state->d;
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
