// extern "C" boundary of libgproshan_ptp.  Every function catches all C++
// exceptions and maps them to a status code; nothing unwinds across the ABI.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <thread>

#include "gp_internal.hpp"

struct gp_che {
	gp::Che m;
};

#define GP_GUARD(...)                                                                  \
	try {                                                                              \
		gp::clear_error();                                                             \
		__VA_ARGS__                                                                    \
	} catch (const std::bad_alloc &) {                                                 \
		return gp::fail(GP_ERR_NOMEM, "host allocation failed");                       \
	} catch (const std::exception &e) {                                                \
		return gp::fail(GP_ERR_INVALID, "exception: %s", e.what());                    \
	} catch (...) {                                                                    \
		return gp::fail(GP_ERR_INVALID, "unknown exception");                          \
	}

static int plan_create_any(gp_che *mesh, const uint32_t *sources, uint32_t n_sources, const double *dhat,
                           const gp_ptp_opts &o, gp_plan **out);

extern "C" {

void gp_ptp_opts_default(gp_ptp_opts *o)
{
	if (!o) return;
	std::memset(o, 0, sizeof(*o));
	o->precision = GP_PREC_F32;
	o->result_buffer = GP_RESULT_LAST_WRITTEN;
	o->schedule = GP_SCHED_RINGS;
	o->coalesce = 1;
	o->tol = GP_PTP_TOL;
	o->launch_mode = GP_LAUNCH_GRAPH;
	o->band_width = 0.0;
	o->local_iters = 8;
	o->band_check = 1;
}

const char *gp_last_error(void) { return gp::last_error(); }

int gp_device_check(char *name_out, size_t cap)
{
	GP_GUARD(return gp::device_check(name_out, cap);)
}

int gp_che_create(const double *xyz, uint32_t n_vertices, const uint32_t *faces, uint32_t n_faces, gp_che **out)
{
	GP_GUARD(
		if (!out) return gp::fail(GP_ERR_INVALID, "out is null");
		*out = nullptr;
		auto c = std::make_unique<gp_che>();
		int rc = gp::build_che(xyz, n_vertices, faces, n_faces, c->m);
		if (rc) return rc;
		*out = c.release();
		return GP_OK;
	)
}

void gp_che_destroy(gp_che *mesh) { delete mesh; }

int gp_che_positions(const gp_che *mesh, double *xyz)
{
	GP_GUARD(
		if (!mesh || !xyz) return gp::fail(GP_ERR_INVALID, "null argument");
		std::memcpy(xyz, mesh->m.gt.data(), sizeof(double) * mesh->m.gt.size());
		return GP_OK;
	)
}

int gp_che_read(const char *path, gp_che **out)
{
	GP_GUARD(
		if (!out) return gp::fail(GP_ERR_INVALID, "out is null");
		*out = nullptr;
		std::vector<double> xyz;
		std::vector<uint32_t> faces;
		int rc = gp::read_mesh(path, xyz, faces);
		if (rc) return rc;
		return gp_che_create(xyz.data(), (uint32_t) (xyz.size() / 3), faces.data(), (uint32_t) (faces.size() / 3), out);
	)
}

int gp_che_sizes(const gp_che *mesh, uint32_t *nv, uint32_t *nf, uint32_t *nhe, uint32_t *ne, int *manifold)
{
	if (!mesh) return gp::fail(GP_ERR_INVALID, "mesh is null");
	if (nv) *nv = mesh->m.nv;
	if (nf) *nf = mesh->m.nf;
	if (nhe) *nhe = mesh->m.nhe;
	if (ne) *ne = mesh->m.ne;
	if (manifold) *manifold = mesh->m.manifold;
	return GP_OK;
}

int gp_che_tables(const gp_che *mesh, uint32_t *vt, uint32_t *ot, uint32_t *evt)
{
	if (!mesh) return gp::fail(GP_ERR_INVALID, "mesh is null");
	const gp::Che &m = mesh->m;
	if (vt && m.nhe) std::memcpy(vt, m.vt.data(), sizeof(uint32_t) * m.nhe);
	if (ot && m.nhe) std::memcpy(ot, m.ot.data(), sizeof(uint32_t) * m.nhe);
	if (evt) std::memcpy(evt, m.evt.data(), sizeof(uint32_t) * m.nv);
	return GP_OK;
}

int gp_che_link(const gp_che *mesh, uint32_t v, uint32_t *he_out, uint32_t cap, uint32_t *n)
{
	if (!mesh || !n) return gp::fail(GP_ERR_INVALID, "null argument");
	const gp::Che &m = mesh->m;
	if (v >= m.nv) return gp::fail(GP_ERR_INVALID, "vertex out of range");
	// Walk the star as che::link does and report half-edge ids.
	uint32_t c = 0;
	const uint32_t stop = m.evt[v];
	for (uint32_t he = stop, guard = 0; he != gp::NIL && guard <= m.nhe; guard++) {
		if (c < cap && he_out) he_out[c] = gp::he_next(he);
		c++;
		const uint32_t pv = gp::he_prev(he);
		if (m.ot[pv] == gp::NIL) {
			if (c < cap && he_out) he_out[c] = pv;
			c++;
		}
		const uint32_t h2 = m.ot[pv];
		he = (h2 != stop) ? h2 : gp::NIL;
	}
	*n = c;
	return c > cap ? gp::fail(GP_ERR_CAPACITY, "link larger than cap") : GP_OK;
}

int gp_che_mean_edge(const gp_che *mesh, double *h)
{
	GP_GUARD(
		if (!mesh || !h) return gp::fail(GP_ERR_INVALID, "null argument");
		*h = mesh->m.mean_edge;
		return GP_OK;
	)
}

int gp_compute_toplesets(const gp_che *mesh, const uint32_t *sources, uint32_t n_sources, uint32_t *toplesets,
                         uint32_t *sorted, uint32_t *limits, uint32_t cap_limits, uint32_t *n_limits)
{
	GP_GUARD(
		if (!mesh || !toplesets || !sorted || !limits || !n_limits || (n_sources && !sources))
			return gp::fail(GP_ERR_INVALID, "null argument");
		std::vector<uint32_t> lim;
		int rc = gp::toplesets_host(mesh->m, sources, n_sources, toplesets, sorted, lim);
		if (rc) return rc;
		*n_limits = (uint32_t) lim.size();
		if (lim.size() > cap_limits) return gp::fail(GP_ERR_CAPACITY, "limits needs %zu entries", lim.size());
		std::memcpy(limits, lim.data(), sizeof(uint32_t) * lim.size());
		return GP_OK;
	)
}

int gp_compute_toplesets_k(const gp_che *mesh, const uint32_t *sources, uint32_t n_sources, uint32_t k,
                           uint32_t *toplesets, uint32_t *sorted, uint32_t *limits, uint32_t cap_limits,
                           uint32_t *n_limits)
{
	GP_GUARD(
		if (!mesh || !toplesets || !sorted || !limits || !n_limits || (n_sources && !sources))
			return gp::fail(GP_ERR_INVALID, "null argument");
		std::vector<uint32_t> lim;
		int rc = gp::toplesets_host(mesh->m, sources, n_sources, toplesets, sorted, lim, k);
		if (rc) return rc;
		*n_limits = (uint32_t) lim.size();
		if (lim.size() > cap_limits) return gp::fail(GP_ERR_CAPACITY, "limits needs %zu entries", lim.size());
		std::memcpy(limits, lim.data(), sizeof(uint32_t) * lim.size());
		return GP_OK;
	)
}

int gp_compute_toplesets_gpu(const gp_che *mesh, const uint32_t *sources, uint32_t n_sources, uint32_t *toplesets,
                             uint32_t *sorted, uint32_t *limits, uint32_t cap_limits, uint32_t *n_limits)
{
	GP_GUARD(
		if (!mesh || !toplesets || !sorted || !limits || !n_limits || (n_sources && !sources))
			return gp::fail(GP_ERR_INVALID, "null argument");
		std::vector<uint32_t> lim;
		// The device tables are cached on the mesh; the mesh itself is not modified.
		int rc = gp::toplesets_gpu(const_cast<gp::Che &>(mesh->m), sources, n_sources, toplesets, sorted, lim);
		if (rc) return rc;
		*n_limits = (uint32_t) lim.size();
		if (lim.size() > cap_limits) return gp::fail(GP_ERR_CAPACITY, "limits needs %zu entries", lim.size());
		std::memcpy(limits, lim.data(), sizeof(uint32_t) * lim.size());
		return GP_OK;
	)
}

static int ptp_gpu_driver(gp_che *mesh, const uint32_t *sources, uint32_t ns, const uint32_t *limits, uint32_t nl,
                          const uint32_t *sorted, double *dist, uint32_t *clusters, double *seconds, int coalesce)
{
	if (!mesh || !dist) return gp::fail(GP_ERR_INVALID, "null argument");
	// The reference's default build computes in FP64 (include/config.h:7-8):
	// the drivers named after its functions return its bits.
	gp_ptp_opts o;
	gp_ptp_opts_default(&o);
	o.precision = GP_PREC_F64;
	o.coalesce = coalesce;
	o.clusters = clusters ? 1 : 0;
	gp_plan *p = nullptr;
	int rc = gp::plan_create_rings(&mesh->m, sources, ns, limits, nl, sorted, o, &p);
	if (rc) return rc;
	std::unique_ptr<gp_plan, void (*)(gp_plan *)> guard(p, gp_plan_destroy);
	gp_ptp_stats st{};
	rc = gp::plan_run(p, &st);
	if (rc) return rc;
	rc = gp::plan_result(p, dist);
	if (!rc && clusters) rc = gp::plan_clusters(p, clusters);
	if (seconds) *seconds = st.kernel_ms * 1e-3;
	return rc;
}

int gp_parallel_toplesets_propagation_gpu(gp_che *mesh, const uint32_t *sources, uint32_t n_sources,
                                          const uint32_t *limits, uint32_t n_limits, const uint32_t *sorted,
                                          double *dist, uint32_t *clusters, double *seconds)
{
	GP_GUARD(return ptp_gpu_driver(mesh, sources, n_sources, limits, n_limits, sorted, dist, clusters, seconds, 0);)
}

int gp_parallel_toplesets_propagation_coalescence_gpu(gp_che *mesh, const uint32_t *sources, uint32_t n_sources,
                                                      const uint32_t *limits, uint32_t n_limits,
                                                      const uint32_t *sorted, double *dist, uint32_t *clusters,
                                                      double *seconds)
{
	GP_GUARD(return ptp_gpu_driver(mesh, sources, n_sources, limits, n_limits, sorted, dist, clusters, seconds, 1);)
}

int gp_parallel_toplesets_propagation_gpu_noalloc(gp_che *mesh, const uint32_t *d_sources, uint32_t n_sources,
                                                  const uint32_t *d_limits, uint32_t n_limits,
                                                  const uint32_t *d_sorted, double *d_dist, void *stream)
{
	GP_GUARD(
		if (!mesh) return gp::fail(GP_ERR_INVALID, "null argument");
		return gp::ptp_gpu_noalloc(mesh->m, d_sources, n_sources, d_limits, n_limits, d_sorted, d_dist, stream);
	)
}

int gp_parallel_toplesets_propagation_cpu(const gp_che *mesh, const uint32_t *sources, uint32_t n_sources,
                                          const uint32_t *limits, uint32_t n_limits, const uint32_t *sorted,
                                          double *dist)
{
	GP_GUARD(
		if (!mesh || !sources || !limits || !sorted || !dist) return gp::fail(GP_ERR_INVALID, "null argument");
		if (n_limits < 2) return gp::fail(GP_ERR_INVALID, "empty limits");
		return gp::ptp_host(mesh->m, sources, n_sources, limits, n_limits, sorted, false, GP_PTP_TOL, dist, nullptr);
	)
}

int gp_parallel_toplesets_propagation_coalescence_cpu(const gp_che *mesh, const uint32_t *sources,
                                                      uint32_t n_sources, const uint32_t *limits,
                                                      uint32_t n_limits, const uint32_t *sorted, double *dist)
{
	GP_GUARD(
		if (!mesh || !sources || !limits || !sorted || !dist) return gp::fail(GP_ERR_INVALID, "null argument");
		if (n_limits < 2) return gp::fail(GP_ERR_INVALID, "empty limits");
		const gp::Che &m = mesh->m;
		gp::Che cm;
		std::vector<uint32_t> inv;
		int rc = gp::coalesce(m, sorted, limits[n_limits - 1], cm, inv);
		if (rc) return rc;
		std::vector<uint32_t> src(n_sources);
		for (uint32_t q = 0; q < n_sources; q++) {
			if (sources[q] >= m.nv || inv[sources[q]] == gp::NIL) return gp::fail(GP_ERR_INVALID, "source not reached");
			src[q] = inv[sources[q]];
		}
		std::vector<double> d(cm.nv);
		rc = gp::ptp_host(cm, src.data(), n_sources, limits, n_limits, nullptr, false, GP_PTP_TOL, d.data(), nullptr);
		if (rc) return rc;
		for (uint32_t v = 0; v < m.nv; v++) dist[v] = inv[v] != gp::NIL ? d[inv[v]] : INFINITY;
		return GP_OK;
	)
}

int gp_normalize_ptp(double *dist, uint32_t n)
{
	if (!dist) return gp::fail(GP_ERR_INVALID, "null argument");
	double mx = 0;
	for (uint32_t v = 0; v < n; v++)
		if (dist[v] < INFINITY && dist[v] > mx) mx = dist[v];
	for (uint32_t v = 0; v < n; v++) dist[v] /= mx;
	return GP_OK;
}

int gp_ptp(gp_che *mesh, const uint32_t *sources, uint32_t n_sources, const gp_ptp_opts *opts, double *dist,
           uint32_t *sorted_out, gp_ptp_stats *stats)
{
	GP_GUARD(
		if (!mesh || !dist || !sources) return gp::fail(GP_ERR_INVALID, "null argument");
		gp_ptp_opts o;
		if (opts) o = *opts; else gp_ptp_opts_default(&o);
		const gp::Che &m = mesh->m;
		// The BFS order the caller asked for (the reference's sorted_index)
		// is built on a host thread while the solve runs.
		int rc0 = GP_OK;
		std::string err0;
		std::thread bfs;
		if (sorted_out) {
			bfs = std::thread([&] {
				try {
					std::vector<uint32_t> lev(m.nv), lim;
					rc0 = gp::toplesets_host(m, sources, n_sources, lev.data(), sorted_out, lim);
					if (rc0) err0 = gp::last_error();
				} catch (...) {
					rc0 = GP_ERR_NOMEM;
					err0 = "toplesets: allocation failed";
				}
			});
		}
		struct Join {
			std::thread &t;
			~Join() { if (t.joinable()) t.join(); }
		} join{ bfs };
		gp_plan *p = nullptr;
		int rc = plan_create_any(mesh, sources, n_sources, nullptr, o, &p);
		if (rc) return rc;
		std::unique_ptr<gp_plan, void (*)(gp_plan *)> guard(p, gp_plan_destroy);
		rc = gp::plan_run(p, stats);
		if (rc) return rc;
		rc = gp::plan_result(p, dist);
		if (rc) return rc;
		if (bfs.joinable()) bfs.join();
		if (rc0) return gp::fail(rc0, "%s", err0.c_str());
		return GP_OK;
	)
}

}  // extern "C"

// Plan of any schedule: rings, or bands from dhat (caller-supplied for the
// learned schedule; Euclidean distance or ring index otherwise).
static int plan_create_any(gp_che *mesh, const uint32_t *sources, uint32_t n_sources, const double *dhat,
                           const gp_ptp_opts &o, gp_plan **out)
{
	const gp::Che &m = mesh->m;
	if (o.fp32_geometry != GP_GEO_AUTO && o.fp32_geometry != GP_GEO_FP32 && o.fp32_geometry != GP_GEO_FP64)
		return gp::fail(GP_ERR_INVALID, "unknown fp32_geometry %d", o.fp32_geometry);
	if (o.schedule == GP_SCHED_PATCHES) {
		gp_ptp_opts op = o;
		if (!(op.band_width > 0)) op.band_width = 128.0;
		return gp::plan_create_patches(&mesh->m, sources, n_sources, dhat, op, out);
	}
	if (o.schedule != GP_SCHED_RINGS && o.schedule != GP_SCHED_LEARNED_BAND && o.schedule != GP_SCHED_EUCLID_BAND &&
	    o.schedule != GP_SCHED_RING_BAND)
		return gp::fail(GP_ERR_INVALID, "unknown schedule %d", o.schedule);
	// Device-built plans read the mesh's own fans.  The reference's coalescence
	// rebuilds the half-edge mesh from the reached faces in ring order
	// (src/geodesics_ptp.cpp:15-38), which on a non-manifold mesh can change a
	// vertex's star (a bowtie vertex whose other wing is unreached becomes an
	// ordinary one); those meshes take the host build, which rebuilds it too.
	const bool dev_ok = m.manifold;
	if (dev_ok && o.device_setup && o.schedule == GP_SCHED_RINGS && o.coalesce && !o.clusters) {
		const int rd = gp::plan_create_device(&mesh->m, sources, n_sources, o, out);
		if (rd != GP_ERR_CAPACITY) return rd;
		gp::clear_error();   // a fan too long for the device build: build on the host
	}
	std::vector<double> d;
	if (!dhat && o.schedule == GP_SCHED_EUCLID_BAND) {
		d.assign(m.nv, INFINITY);
		#pragma omp parallel for schedule(static)
		for (int64_t v = 0; v < (int64_t) m.nv; v++) {
			double best = INFINITY;
			for (uint32_t q = 0; q < n_sources; q++) {
				const double *a = &m.gt[3 * (size_t) v], *b = &m.gt[3 * (size_t) sources[q]];
				best = std::min(best, std::sqrt((a[0] - b[0]) * (a[0] - b[0]) + (a[1] - b[1]) * (a[1] - b[1]) +
				                                (a[2] - b[2]) * (a[2] - b[2])));
			}
			d[v] = best;
		}
		dhat = d.data();
	}
	// Bands of predicted distance take the place of the rings in the
	// reference's uncapped window loop; the plan keeps the band schedule, which
	// adds the check sweeps over every vertex at the end of a run.
	const double width = (o.band_width > 0 ? o.band_width : 1.0) * m.mean_edge;
	gp_ptp_opts ob = o;
	ob.coalesce = 1;
	ob.uncapped = 1;
	if (dev_ok && o.device_setup && !o.clusters && dhat &&
	    (o.schedule == GP_SCHED_LEARNED_BAND || o.schedule == GP_SCHED_EUCLID_BAND)) {
		const int rd = gp::plan_create_device(&mesh->m, sources, n_sources, ob, out, dhat, width);
		if (rd != GP_ERR_CAPACITY) return rd;
		gp::clear_error();   // a fan too long or a band index too large: build on the host
	}
	std::vector<uint32_t> lev(m.nv), sorted(m.nv), lim;
	int rc = gp::toplesets_host(m, sources, n_sources, lev.data(), sorted.data(), lim);
	if (rc) return rc;
	if (o.schedule == GP_SCHED_RINGS)
		return gp::plan_create_rings(&mesh->m, sources, n_sources, lim.data(), (uint32_t) lim.size(), sorted.data(), o, out);
	if (!dhat) {
		if (o.schedule != GP_SCHED_RING_BAND)
			return gp::fail(GP_ERR_INVALID, "learned bands need predicted distances (gp_plan_create_bands)");
		d.assign(m.nv, INFINITY);
		for (uint32_t v = 0; v < m.nv; v++) d[v] = lev[v] == gp::NIL ? INFINITY : (double) lev[v] * m.mean_edge;
		dhat = d.data();
	}
	// Host-built bands: the vertices of a band in Morton order (the coalesced
	// numbering of the plan).
	std::vector<uint32_t> order, blim;
	rc = gp::band_order(m, sources, n_sources, lev.data(), dhat, width, order, blim);
	if (rc) return rc;
	return gp::plan_create_rings(&mesh->m, sources, n_sources, blim.data(), (uint32_t) blim.size(), order.data(), ob, out);
}

extern "C" {

int gp_plan_create(gp_che *mesh, const uint32_t *sources, uint32_t n_sources, const gp_ptp_opts *opts, gp_plan **out)
{
	GP_GUARD(
		if (!mesh || !sources || !out) return gp::fail(GP_ERR_INVALID, "null argument");
		gp_ptp_opts o;
		if (opts) o = *opts; else gp_ptp_opts_default(&o);
		return plan_create_any(mesh, sources, n_sources, nullptr, o, out);
	)
}

int gp_plan_create_bands(gp_che *mesh, const uint32_t *sources, uint32_t n_sources, const double *dhat,
                         const gp_ptp_opts *opts, gp_plan **out)
{
	GP_GUARD(
		if (!mesh || !sources || !out || !dhat) return gp::fail(GP_ERR_INVALID, "null argument");
		gp_ptp_opts o;
		if (opts) o = *opts; else gp_ptp_opts_default(&o);
		if (o.schedule == GP_SCHED_RINGS) return gp::fail(GP_ERR_INVALID, "gp_plan_create_bands needs a band schedule");
		return plan_create_any(mesh, sources, n_sources, dhat, o, out);
	)
}

int gp_plan_run(gp_plan *plan, gp_ptp_stats *stats) { GP_GUARD(return gp::plan_run(plan, stats);) }
int gp_plan_autotune(gp_plan *plan, gp_tune *out) { GP_GUARD(return gp::plan_autotune(plan, out);) }

int gp_plan_create_bands_tuned(gp_che *mesh, const uint32_t *sources, uint32_t n_sources, const double *dhat,
                               const gp_ptp_opts *opts, const double *widths, uint32_t n_widths, gp_plan **out,
                               double *width_out)
{
	GP_GUARD(
		if (!mesh || !sources || !out || !dhat) return gp::fail(GP_ERR_INVALID, "null argument");
		*out = nullptr;
		gp_ptp_opts o;
		if (opts) o = *opts; else gp_ptp_opts_default(&o);
		if (o.schedule == GP_SCHED_RINGS || o.schedule == GP_SCHED_PATCHES)
			return gp::fail(GP_ERR_INVALID, "gp_plan_create_bands_tuned needs a band schedule");
		static const double defaults[] = { 0.45, 0.5, 0.6, 0.75, 1.0 };
		// With the default candidates, a width chosen before for this mesh
		// shape is taken without timing, and a new choice is remembered.
		const bool use_cache = !widths || !n_widths;
		if (use_cache) { widths = defaults; n_widths = 5; }
		double cached = 0;
		if (use_cache && gp::tune_width_lookup(mesh->m, o.schedule, o.precision, &cached)) {
			widths = &cached;
			n_widths = 1;
		}
		std::unique_ptr<gp_plan, void (*)(gp_plan *)> best(nullptr, gp_plan_destroy);
		double best_ms = 1e300, best_w = 0;
		for (uint32_t q = 0; q < n_widths; q++) {
			if (!(widths[q] > 0)) return gp::fail(GP_ERR_INVALID, "band width %g must be positive", widths[q]);
			o.band_width = widths[q];
			gp_plan *p = nullptr;
			int rc = plan_create_any(mesh, sources, n_sources, dhat, o, &p);
			if (rc) return rc;
			std::unique_ptr<gp_plan, void (*)(gp_plan *)> cand(p, gp_plan_destroy);
			gp_tune t{};
			if ((rc = gp::plan_autotune(p, &t))) return rc;
			// A cached shape was not timed here: time the plan with it.
			if (t.candidates == 0) {
				gp_ptp_stats st{};
				double ms = 0;
				for (int k = 0; k < 3 && !rc; k++) {
					rc = gp::plan_run(p, &st);
					if (k) ms += st.kernel_ms / 2;
				}
				if (rc) return rc;
				t.kernel_ms = ms;
			}
			if (t.kernel_ms < best_ms) { best_ms = t.kernel_ms; best_w = widths[q]; best = std::move(cand); }
		}
		if (use_cache && n_widths > 1) gp::tune_width_store(mesh->m, o.schedule, o.precision, best_w);
		if (width_out) *width_out = best_w;
		*out = best.release();
		return GP_OK;
	)
}
int gp_plan_clusters(gp_plan *plan, uint32_t *clusters) { GP_GUARD(return gp::plan_clusters(plan, clusters);) }
int gp_plan_result(gp_plan *plan, double *dist) { GP_GUARD(return gp::plan_result(plan, dist);) }
int gp_plan_trace(gp_plan *plan, uint32_t *trace_out, uint32_t cap, uint32_t *n)
{
	GP_GUARD(return gp::plan_trace(plan, trace_out, cap, n);)
}
void gp_plan_destroy(gp_plan *plan);

int gp_mlp_create(uint32_t seed, gp_mlp **out)
{
	GP_GUARD(
		if (!out) return gp::fail(GP_ERR_INVALID, "null argument");
		return gp::mlp_create(seed, out);
	)
}

int gp_mlp_params(gp_mlp *mlp, float *get, const float *set, uint32_t n)
{
	GP_GUARD(
		if (!mlp) return gp::fail(GP_ERR_INVALID, "null argument");
		return gp::mlp_params(mlp, get, set, n);
	)
}

int gp_mlp_features(gp_che *mesh, const uint32_t *sources, uint32_t n_sources, float *features, float *euclid)
{
	GP_GUARD(
		if (!mesh || !sources || !n_sources) return gp::fail(GP_ERR_INVALID, "null argument");
		std::vector<uint32_t> lev(mesh->m.nv), srt(mesh->m.nv), lim;
		int rc = gp::toplesets_host(mesh->m, sources, n_sources, lev.data(), srt.data(), lim);
		if (rc) return rc;
		return gp::mlp_features(mesh->m, sources, n_sources, lev.data(), features, euclid);
	)
}

int gp_mlp_forward(gp_mlp *mlp, const float *features, uint32_t n, float *y)
{
	GP_GUARD(
		if (!mlp || !features || !y || !n) return gp::fail(GP_ERR_INVALID, "null argument");
		return gp::mlp_forward(mlp, features, n, y);
	)
}

int gp_mlp_train(gp_mlp *mlp, const float *features, const float *target, uint32_t n, uint32_t steps,
                 uint32_t batch, float lr, uint32_t seed, float *grad, float *loss)
{
	GP_GUARD(
		if (!mlp || !features || !target) return gp::fail(GP_ERR_INVALID, "null argument");
		return gp::mlp_train(mlp, features, target, n, steps, batch, lr, seed, grad, loss);
	)
}

int gp_mlp_predict(gp_mlp *mlp, gp_che *mesh, const uint32_t *sources, uint32_t n_sources, double *dhat, float *ms)
{
	GP_GUARD(
		if (!mlp || !mesh || !sources || !n_sources || !dhat) return gp::fail(GP_ERR_INVALID, "null argument");
		// The ring-index feature: BFS levels of the sources, on the device.
		std::vector<uint32_t> lev(mesh->m.nv);
		int rc = gp::levels_gpu(mesh->m, sources, n_sources, lev.data());
		if (rc) return rc;
		return gp::mlp_predict(mlp, mesh->m, sources, n_sources, lev.data(), dhat, ms);
	)
}

int gp_plan_create_learned(gp_che *mesh, const uint32_t *sources, uint32_t n_sources, gp_mlp *mlp,
                           const gp_ptp_opts *opts, gp_plan **out)
{
	GP_GUARD(
		if (!mlp || !mesh || !sources || !n_sources || !out) return gp::fail(GP_ERR_INVALID, "null argument");
		gp_ptp_opts o;
		if (opts) o = *opts; else gp_ptp_opts_default(&o);
		o.schedule = GP_SCHED_LEARNED_BAND;
		if (o.fp32_geometry != GP_GEO_AUTO && o.fp32_geometry != GP_GEO_FP32 && o.fp32_geometry != GP_GEO_FP64)
			return gp::fail(GP_ERR_INVALID, "unknown fp32_geometry %d", o.fp32_geometry);
		if (o.device_setup && !o.clusters && mesh->m.manifold) {
			// Everything on the device: one BFS serves the features' ring
			// index and the plan's reached set; dhat never leaves the GPU.
			gp_ptp_opts ob = o;
			ob.coalesce = 1;
			ob.uncapped = 1;
			const double width = (o.band_width > 0 ? o.band_width : 1.0) * mesh->m.mean_edge;
			const int rd = gp::plan_create_device(&mesh->m, sources, n_sources, ob, out, nullptr, width, mlp);
			if (rd != GP_ERR_CAPACITY) return rd;
			gp::clear_error();   // a fan too long or a band index too large: build on the host
		}
		std::vector<double> dhat(mesh->m.nv);
		int rc = gp_mlp_predict(mlp, mesh, sources, n_sources, dhat.data(), nullptr);
		if (rc) return rc;
		return plan_create_any(mesh, sources, n_sources, dhat.data(), o, out);
	)
}

int gp_batch_create(gp_che *const *meshes, uint32_t n_meshes, const uint32_t *job_mesh, const uint32_t *job_source,
                    uint32_t n_jobs, double tol, uint32_t block, gp_batch **out)
{
	GP_GUARD(
		if (!meshes || !out) return gp::fail(GP_ERR_INVALID, "null argument");
		*out = nullptr;
		std::vector<gp::Che *> ms(n_meshes);
		for (uint32_t k = 0; k < n_meshes; k++) {
			if (!meshes[k]) return gp::fail(GP_ERR_INVALID, "batch: mesh %u is null", k);
			ms[k] = &meshes[k]->m;
		}
		return gp::batch_create(ms.data(), n_meshes, job_mesh, job_source, n_jobs, tol, block, out);
	)
}

int gp_batch_run(gp_batch *batch, gp_batch_stats *stats)
{
	GP_GUARD(return gp::batch_run(batch, stats);)
}

int gp_batch_result(const gp_batch *batch, uint32_t job, double *dist, uint32_t cap, uint32_t *sweeps)
{
	GP_GUARD(return gp::batch_result(batch, job, dist, cap, sweeps);)
}

int gp_farthest_point_sampling_ptp_gpu(gp_che *mesh, uint32_t *samples, uint32_t n_init, uint32_t cap,
                                       uint64_t n, double radio, int precision, uint32_t *n_out,
                                       double *radio_out, double *seconds)
{
	GP_GUARD(
		if (!mesh || !samples) return gp::fail(GP_ERR_INVALID, "null argument");
		if (precision != GP_PREC_F32 && precision != GP_PREC_F64) return gp::fail(GP_ERR_INVALID, "bad precision");
		return gp::fps_gpu(mesh->m, samples, n_init, cap, n, radio, precision == GP_PREC_F64, n_out, radio_out,
		                   seconds);
	)
}

int gp_ptp_iter_error(gp_che *mesh, const uint32_t *sources, uint32_t n_sources, const double *exact,
                      int precision, uint32_t *iter, double *err, uint32_t cap, uint32_t *n_err,
                      uint32_t *band, uint32_t cap_band, uint32_t *n_band)
{
	GP_GUARD(
		if (!mesh) return gp::fail(GP_ERR_INVALID, "null argument");
		if (precision != GP_PREC_F32 && precision != GP_PREC_F64) return gp::fail(GP_ERR_INVALID, "bad precision");
		return gp::ptp_iter_error(&mesh->m, sources, n_sources, exact, precision == GP_PREC_F64, iter, err, cap, n_err,
		                          band, cap_band, n_band);
	)
}

}  // extern "C"
