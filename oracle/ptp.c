/*
 * TEST INFRASTRUCTURE -- NOT PART OF THE PRODUCT.
 * Oracle: update_step and the PTP window schedule in FP64 (reference default
 * build) and FP32 (reference SINGLE_P build), the converged fixed point,
 * compute_error and normalize_ptp.
 *
 * Restates src/geodesics_ptp.cpp:15-276, src/cuda/geodesics_ptp.cu:174-232,
 * src/cuda/geodesics_ptp_coalescence.cu:102-160,
 * src/cuda/test_geodesics_ptp_coalescence.cu:214-266 and
 * src/test_geodesics_ptp.cpp:362-371.
 */
#include "oracle.h"

#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#define REAL double
#define SFX(name) name##_f64
#define SQRT sqrt
#define FABS fabs
#include "ptp_impl.h"
#undef REAL
#undef SFX
#undef SQRT
#undef FABS

#define REAL float
#define SFX(name) name##_f32
#define SQRT sqrtf
#define FABS fabsf
#include "ptp_impl.h"
#undef REAL
#undef SFX
#undef SQRT
#undef FABS

double or_update_step(const or_che *m, const double *dist, uint32_t he)
{
	mview_f64 mv = { m->nv, m->gt, m->vt, m->ot, m->evt };
	return update_step_f64(&mv, dist, he);
}

float or_update_step_f32(const float *gt, const uint32_t *vt, const float *dist, uint32_t he)
{
	mview_f32 mv = { 0, gt, vt, NULL, NULL };
	return update_step_f32(&mv, dist, he);
}

int or_ptp(const or_che *m, const uint32_t *src, uint32_t ns,
           const uint32_t *limits, uint32_t nlimits, const uint32_t *sorted,
           int variant, double tol, int n_threads,
           double *dist, or_sweep *trace, uint32_t trace_cap, uint32_t *n_sweeps)
{
	return ptp_f64(m, src, ns, limits, nlimits, sorted, variant, tol, n_threads,
	               dist, trace, trace_cap, n_sweeps, NULL);
}

int or_ptp_clusters(const or_che *m, const uint32_t *src, uint32_t ns,
                    const uint32_t *limits, uint32_t nlimits, const uint32_t *sorted,
                    int variant, double tol, int n_threads,
                    double *dist, uint32_t *clusters, uint32_t *n_sweeps)
{
	if (!clusters) return OR_EINVAL;
	return ptp_f64(m, src, ns, limits, nlimits, sorted, variant, tol, n_threads,
	               dist, NULL, 0, n_sweeps, clusters);
}

int or_ptp_f32(const or_che *m, const uint32_t *src, uint32_t ns,
               const uint32_t *limits, uint32_t nlimits, const uint32_t *sorted,
               int variant, double tol, int n_threads,
               float *dist, or_sweep *trace, uint32_t trace_cap, uint32_t *n_sweeps)
{
	return ptp_f32(m, src, ns, limits, nlimits, sorted, variant, tol, n_threads,
	               dist, trace, trace_cap, n_sweeps, NULL);
}

/* FIFO label-correcting relaxation to the fixed point (see oracle.h).
 * Pins: on flat meshes the fixed point is within the planar-update
 * discretisation error of the Euclidean distance and below the reference
 * schedule's result everywhere (tests/test_oracle_ptp.py). */
int or_ptp_fixed_point(const or_che *m, const uint32_t *src, uint32_t ns,
                       double rel_eps, double *dist, uint64_t *n_relax)
{
	const uint32_t nv = m->nv;
	mview_f64 mv = { nv, m->gt, m->vt, m->ot, m->evt };
	uint8_t *flag = (uint8_t *) calloc(nv, 1);      /* bit0 queued, bit1 source */
	uint32_t *q = (uint32_t *) malloc(sizeof(uint32_t) * ((size_t) nv + 1));
	if (!flag || !q) { free(flag); free(q); return OR_ENOMEM; }
	for (uint32_t v = 0; v < nv; v++) dist[v] = INFINITY;
	for (uint32_t s = 0; s < ns; s++) {
		if (src[s] >= nv) { free(flag); free(q); return OR_EINVAL; }
		dist[src[s]] = 0;
		flag[src[s]] |= 2;
	}
	const size_t qn = (size_t) nv + 1;
	size_t head = 0, tail = 0;
	uint32_t lk[64];
	for (uint32_t s = 0; s < ns; s++) {
		const uint32_t n = or_link(m, src[s], lk, 64);
		for (uint32_t k = 0; k < n && k < 64; k++) {
			const uint32_t u = m->vt[lk[k]];
			if (!flag[u]) { flag[u] |= 1; q[tail] = u; tail = (tail + 1) % qn; }
		}
	}
	uint64_t relax = 0;
	while (head != tail) {
		const uint32_t v = q[head]; head = (head + 1) % qn;
		flag[v] &= (uint8_t) ~1u;
		const double nd = relax_f64(&mv, dist, v);
		relax++;
		const double od = dist[v];
		if (nd < od && (od == INFINITY || od - nd > rel_eps * od)) {
			dist[v] = nd;
			const uint32_t n = or_link(m, v, lk, 64);
			for (uint32_t k = 0; k < n && k < 64; k++) {
				const uint32_t u = m->vt[lk[k]];
				if (!flag[u]) { flag[u] |= 1; q[tail] = u; tail = (tail + 1) % qn; }
			}
		}
	}
	if (n_relax) *n_relax = relax;
	free(flag); free(q);
	return OR_OK;
}

int or_fps(const or_che *m, uint32_t *samples, uint32_t n_init, uint32_t cap, uint64_t n, double radio,
           int f32, uint32_t *n_out, double *radio_out)
{
	const uint32_t nv = m->nv;
	*n_out = n_init;
	*radio_out = INFINITY;
	if (n_init == 0) return OR_EINVAL;
	if (n >= nv) n = nv >> 1;                                  /* :123 */
	const uint64_t rounds = n > n_init ? n - n_init : 0;      /* :125 (without the size_t wrap) */
	if (n_init + rounds > cap) return OR_ECAPACITY;
	uint32_t *levels = (uint32_t *) malloc(sizeof(uint32_t) * nv);
	uint32_t *sorted = (uint32_t *) malloc(sizeof(uint32_t) * nv);
	uint32_t *limits = (uint32_t *) malloc(sizeof(uint32_t) * ((size_t) nv + 2));
	double *d64 = (double *) malloc(sizeof(double) * nv);
	float *d32 = (float *) malloc(sizeof(float) * nv);
	int rc = OR_OK;
	if (!levels || !sorted || !limits || !d64 || !d32) { rc = OR_ENOMEM; goto out; }
	uint32_t count = n_init;
	double max_dist = INFINITY;
	for (uint64_t left = rounds; left-- > 0 && max_dist > radio;) {     /* :132 */
		uint32_t nl = 0, nsw = 0;
		rc = or_toplesets(m, samples, count, OR_NIL, levels, sorted, limits, nv + 2, &nl);
		if (rc) goto out;
		if (f32) rc = or_ptp_f32(m, samples, count, limits, nl, sorted, 2 /* GPU_PLAIN */, 1e-3, 0, d32, NULL, 0, &nsw);
		else rc = or_ptp(m, samples, count, limits, nl, sorted, 2 /* GPU_PLAIN */, 1e-3, 0, d64, NULL, 0, &nsw);
		if (rc) goto out;
		/* First index of the largest |value| (:140-142). */
		uint32_t f = 0;
		double best = -1;
		for (uint32_t v = 0; v < nv; v++) {
			const double x = f32 ? fabs((double) d32[v]) : fabs(d64[v]);
			if (x > best) { best = x; f = v; }
		}
		if (best == 0) break;
		if (radio > 0 || left == 0) max_dist = best;                   /* :144-145 */
		samples[count++] = f;                                          /* :147 */
	}
	*n_out = count;
	*radio_out = max_dist;
out:
	free(levels); free(sorted); free(limits); free(d64); free(d32);
	return rc;
}

int or_iter_error(const or_che *m, const uint32_t *src, uint32_t ns, const double *exact, int f32,
                  uint32_t *iter, double *err, uint32_t cap, uint32_t *n_err,
                  uint32_t *band, uint32_t cap_band, uint32_t *n_band)
{
	const uint32_t nv = m->nv;
	*n_err = *n_band = 0;
	uint32_t *levels = (uint32_t *) malloc(sizeof(uint32_t) * nv);
	uint32_t *sorted = (uint32_t *) malloc(sizeof(uint32_t) * nv);
	uint32_t *limits = (uint32_t *) malloc(sizeof(uint32_t) * ((size_t) nv + 2));
	uint32_t *inv = (uint32_t *) malloc(sizeof(uint32_t) * nv);
	uint32_t *srcc = (uint32_t *) malloc(sizeof(uint32_t) * (ns ? ns : 1));
	double *exc = (double *) malloc(sizeof(double) * nv);
	double *V = (double *) malloc(sizeof(double) * 3 * (size_t) nv);
	uint32_t *F = (uint32_t *) malloc(sizeof(uint32_t) * (m->nhe ? m->nhe : 1));
	float *gt32 = NULL;
	or_che cm;
	int built = 0, rc = OR_OK;
	uint32_t nl = 0;
	if (!levels || !sorted || !limits || !inv || !srcc || !exc || !V || !F) { rc = OR_ENOMEM; goto out; }
	rc = or_toplesets(m, src, ns, OR_NIL, levels, sorted, limits, nv + 2, &nl);
	if (rc) goto out;
	if (limits[nl - 1] != nv) { rc = OR_EINVAL; goto out; }   /* the reference renumbers every vertex */
	/* Coalesced copy (src/cuda/test_geodesics_ptp_coalescence.cu:27-44). */
	for (uint32_t q = 0; q < nv; q++) {
		for (int k = 0; k < 3; k++) V[3 * (size_t) q + k] = m->gt[3 * (size_t) sorted[q] + k];
		inv[sorted[q]] = q;
		exc[q] = exact[sorted[q]];
	}
	for (uint32_t he = 0; he < m->nhe; he++) F[he] = inv[m->vt[he]];
	rc = or_che_build(V, nv, F, m->nhe / 3, &cm);
	if (rc) goto out;
	built = 1;
	for (uint32_t q = 0; q < ns; q++) srcc[q] = inv[src[q]];
	if (f32) {
		gt32 = (float *) malloc(sizeof(float) * 3 * (size_t) nv);
		if (!gt32) { rc = OR_ENOMEM; goto out; }
		for (size_t q = 0; q < 3 * (size_t) nv; q++) gt32[q] = (float) cm.gt[q];
		mview_f32 mv = { nv, gt32, cm.vt, cm.ot, cm.evt };
		rc = iter_error_f32(&mv, srcc, ns, limits, nl, exc, 1e-3, iter, err, cap, n_err, band, cap_band, n_band);
	} else {
		mview_f64 mv = { nv, cm.gt, cm.vt, cm.ot, cm.evt };
		rc = iter_error_f64(&mv, srcc, ns, limits, nl, exc, 1e-3, iter, err, cap, n_err, band, cap_band, n_band);
	}
out:
	if (built) or_che_free(&cm);
	free(levels); free(sorted); free(limits); free(inv); free(srcc); free(exc); free(V); free(F); free(gt32);
	return rc;
}

double or_compute_error(const double *dist, const double *exact, uint32_t n, uint32_t s)
{
	double e = 0;
	#pragma omp parallel for reduction(+: e)
	for (int64_t v = 0; v < (int64_t) n; v++)
		if (exact[v] > 0) e += fabs(dist[v] - exact[v]) / exact[v];
	return e * 100 / (double) (n - s);
}

void or_normalize_ptp(double *dist, uint32_t n)
{
	double mx = 0;
	#pragma omp parallel for reduction(max: mx)
	for (int64_t v = 0; v < (int64_t) n; v++)
		if (dist[v] < INFINITY && dist[v] > mx) mx = dist[v];
	#pragma omp parallel for
	for (int64_t v = 0; v < (int64_t) n; v++)
		dist[v] /= mx;
}
