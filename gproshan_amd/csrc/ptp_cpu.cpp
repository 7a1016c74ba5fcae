// Host (OpenMP, FP64) parallel toplesets propagation: the PTP_CPU option of
// the front end.  It runs the reference's window schedule
// (src/geodesics_ptp.cpp:61-109, 137-189) over the triangle fans of the
// mesh.  Compiled with -ffp-contract=off.
#include "gp_internal.hpp"
#include "update_step.hpp"

#include <omp.h>

#include <cmath>
#include <cstring>

namespace gp {

static inline double relax_host(const Che &m, const double *od, uint32_t v)
{
	double best = od[v];
	const uint32_t s = m.fan_off[v], e = m.fan_off[v + 1];
	if (e - s < 2) return best;
	const double *gv = &m.gt[3 * (size_t) v];
	for (uint32_t k = s; k + 1 < e; k++) {
		const uint32_t a = m.fan[k], b = m.fan[k + 1];
		const double p = update_ref_f64(gv, &m.gt[3 * (size_t) a], &m.gt[3 * (size_t) b], od[a], od[b]);
		if (p < best) best = p;
	}
	return best;
}

int ptp_host(const Che &m, const uint32_t *src, uint32_t ns, const uint32_t *limits, uint32_t nl,
             const uint32_t *order, bool last_written, double tol, double *dist_out, uint32_t *sweeps_out)
{
	const uint32_t n = m.nv;
	std::vector<double> b0(n, INFINITY), b1(n, INFINITY);
	for (uint32_t q = 0; q < ns; q++) {
		if (src[q] >= n) return fail(GP_ERR_INVALID, "ptp: source %u out of range", src[q]);
		b0[src[q]] = b1[src[q]] = 0;
	}
	double *pd[2] = { b0.data(), b1.data() };
	uint32_t d = 0, i = 1, j = 2, iter = 0, sweeps = 0;
	const uint32_t max_iter = nl << 1;
	if (nl >= 3) {
		while (i < j && iter++ < max_iter) {
			if (i < (j >> 1)) i = j >> 1;
			const uint32_t start = limits[i], end = limits[j], cend = limits[i + 1];
			const double *od = pd[d];
			double *nd = pd[!d];
			uint32_t bad = 0;
			#pragma omp parallel for schedule(static) reduction(+: bad)
			for (int64_t k = start; k < (int64_t) end; k++) {
				const uint32_t v = order ? order[k] : (uint32_t) k;
				const double nv_ = relax_host(m, od, v);
				nd[v] = nv_;
				if (k < (int64_t) cend) bad += !(std::fabs(nv_ - od[v]) / od[v] < tol);
			}
			sweeps++;
			if (bad == 0) i++;
			if (j < nl - 1) j++;
			d = !d;
		}
	}
	const double *res = last_written ? pd[d] : pd[!d];
	std::memcpy(dist_out, res, sizeof(double) * n);
	if (sweeps_out) *sweeps_out = sweeps;
	return GP_OK;
}

}  // namespace gp
