/*
 * TEST INFRASTRUCTURE -- NOT PART OF THE PRODUCT.
 * Body of the PTP oracle, included twice by ptp.c: once with REAL = double
 * (the reference's default build) and once with REAL = float (the reference's
 * SINGLE_P build, include/config.h:7-8, include/include.h:21-25).
 *
 * Requires: REAL, SFX(name), SQRT, FABS.
 */

typedef struct {
	uint32_t nv;
	const REAL *gt;         /* 3*nv */
	const uint32_t *vt, *ot, *evt;
} SFX(mview);

/* update_step (src/geodesics_ptp.cpp:201-262).  The planar-wavefront update
 * of vertex VT[he] from its two fan neighbours: solve the 2x2 Gram system of
 * the triangle, take the larger root of the unit-gradient quadratic, and fall
 * back to the shorter edge path when an input is infinite, the discriminant
 * is negative, or the front does not arrive from inside the wedge.  The
 * operations and their order are the reference's; the oracle is built without
 * FMA contraction.
 * Pins (tests/test_oracle_update_step.py): planar waves inside the wedge give
 * n.v + c to rounding; one infinite input gives the other edge path; a front
 * from outside the wedge gives the edge path; an equilateral closed form;
 * det == 0 gives NaN; 50-digit mpmath-free long-double re-evaluation. */
static REAL SFX(update_step)(const SFX(mview) *m, const REAL *dist, uint32_t he)
{
	const uint32_t x0 = m->vt[or_next(he)];
	const uint32_t x1 = m->vt[or_prev(he)];
	const uint32_t x2 = m->vt[he];

	const REAL *g0 = m->gt + 3 * (size_t) x0;
	const REAL *g1 = m->gt + 3 * (size_t) x1;
	const REAL *g2 = m->gt + 3 * (size_t) x2;
	const REAL X0[3] = { g0[0] - g2[0], g0[1] - g2[1], g0[2] - g2[2] };
	const REAL X1[3] = { g1[0] - g2[0], g1[1] - g2[1], g1[2] - g2[2] };

	const REAL t0 = dist[x0];
	const REAL t1 = dist[x1];

	const REAL q00 = X0[0] * X0[0] + X0[1] * X0[1] + X0[2] * X0[2];
	const REAL q01 = X0[0] * X1[0] + X0[1] * X1[1] + X0[2] * X1[2];
	const REAL q10 = X1[0] * X0[0] + X1[1] * X0[1] + X1[2] * X0[2];
	const REAL q11 = X1[0] * X1[0] + X1[1] * X1[1] + X1[2] * X1[2];

	const REAL det = q00 * q11 - q01 * q10;
	const REAL Q00 = q11 / det;
	const REAL Q01 = -q01 / det;
	const REAL Q10 = -q10 / det;
	const REAL Q11 = q00 / det;

	const REAL delta = t0 * (Q00 + Q10) + t1 * (Q01 + Q11);
	const REAL dis = delta * delta -
	                 (Q00 + Q01 + Q10 + Q11) *
	                 (t0 * t0 * Q00 + t0 * t1 * (Q10 + Q01) + t1 * t1 * Q11 - 1);

	REAL p = (delta + SQRT(dis)) / (Q00 + Q01 + Q10 + Q11);

	const REAL tp0 = t0 - p;
	const REAL tp1 = t1 - p;

	const REAL n[3] = {
		tp0 * (X0[0] * Q00 + X1[0] * Q10) + tp1 * (X0[0] * Q01 + X1[0] * Q11),
		tp0 * (X0[1] * Q00 + X1[1] * Q10) + tp1 * (X0[1] * Q01 + X1[1] * Q11),
		tp0 * (X0[2] * Q00 + X1[2] * Q10) + tp1 * (X0[2] * Q01 + X1[2] * Q11) };

	const REAL cond0 = X0[0] * n[0] + X0[1] * n[1] + X0[2] * n[2];
	const REAL cond1 = X1[0] * n[0] + X1[1] * n[1] + X1[2] * n[2];

	const REAL c0 = cond0 * Q00 + cond1 * Q01;
	const REAL c1 = cond0 * Q10 + cond1 * Q11;

	if (t0 == (REAL) INFINITY || t1 == (REAL) INFINITY || dis < 0 || c0 >= 0 || c1 >= 0) {
		const REAL dp0 = dist[x0] + SQRT(X0[0] * X0[0] + X0[1] * X0[1] + X0[2] * X0[2]);
		const REAL dp1 = dist[x1] + SQRT(X1[0] * X1[0] + X1[1] * X1[1] + X1[2] * X1[2]);
		p = (dp1 < dp0) ? dp1 : dp0;
	}
	return p;
}

/* One Jacobi relaxation of vertex v: copy the old value, then keep the
 * smallest update over the star (src/geodesics_ptp.cpp:80-93, 157-170;
 * src/cuda/geodesics_ptp.cu:244-253).  NaN updates never win the '<'. */
static inline REAL SFX(relax)(const SFX(mview) *m, const REAL *old_d, uint32_t v)
{
	REAL best = old_d[v];
	const uint32_t stop = m->evt[v];
	for (uint32_t he = stop; he != OR_NIL; ) {
		const REAL p = SFX(update_step)(m, old_d, he);
		if (p < best) best = p;
		const uint32_t h2 = m->ot[or_prev(he)];
		he = (h2 != stop) ? h2 : OR_NIL;
	}
	return best;
}

/* The same relaxation carrying the cluster (Voronoi) label: when a triangle
 * lowers the value, the vertex takes the label of the nearer (by old
 * distance, ties to the next vertex) of that triangle's two other vertices --
 * the rule of the reference's GPU drivers, read from the old buffers
 * (src/cuda/geodesics_ptp.cu:257-282, src/cuda/geodesics_ptp_coalescence.cu:185-207). */
static inline REAL SFX(relax_cl)(const SFX(mview) *m, const REAL *old_d, const uint32_t *old_c, uint32_t v,
                                 uint32_t *new_c)
{
	REAL best = old_d[v];
	uint32_t c = old_c[v];
	const uint32_t stop = m->evt[v];
	for (uint32_t he = stop; he != OR_NIL; ) {
		const REAL p = SFX(update_step)(m, old_d, he);
		if (p < best) {
			best = p;
			const uint32_t xp = m->vt[or_prev(he)], xn = m->vt[or_next(he)];
			c = old_d[xp] < old_d[xn] ? old_c[xp] : old_c[xn];
		}
		const uint32_t h2 = m->ot[or_prev(he)];
		he = (h2 != stop) ? h2 : OR_NIL;
	}
	*new_c = c;
	return best;
}

/* The window schedule shared by the four reference drivers
 * (src/geodesics_ptp.cpp:61-109 and 137-189; src/cuda/geodesics_ptp.cu:198-231;
 * src/cuda/geodesics_ptp_coalescence.cu:125-159), and, with OR_UNCAPPED, by the
 * iter_error loop (src/cuda/test_geodesics_ptp_coalescence.cu:223-248).
 * order == NULL means the mesh is already numbered by rings (coalescence).
 * Returns the buffer selected by OR_LAST_WRITE.
 * Pins (tests/test_oracle_ptp.py): band traces stepped by hand on a strip
 * mesh; monotone non-increasing distances per sweep; last-read and
 * last-written buffers differ by exactly one sweep; the K/2 rule and the
 * 2*|limits| cap observed in traces; flat meshes approach Euclidean
 * distance as the mesh is refined. */
static REAL *SFX(schedule)(const SFX(mview) *m, const uint32_t *limits, uint32_t nl,
                           const uint32_t *order, int variant, double tol, int nthreads,
                           REAL *pd0, REAL *pd1, REAL *err,
                           or_sweep *trace, uint32_t cap, uint32_t *nsw, uint32_t **pc)
{
	REAL *pd[2] = { pd0, pd1 };
	const int uncapped = (variant & OR_UNCAPPED) != 0;
	uint32_t d = 0, i = 1, j = 2, iter = 0, sweeps = 0;
	const uint32_t max_iter = nl << 1;                     /* :66-67 */

	/* With only the source ring there is no ring 1; the reference would read
	 * limits[2] out of bounds.  No sweep is run. */
	if (nl < 3) { *nsw = 0; return (variant & OR_LAST_WRITE) ? pd[d] : pd[!d]; }

	while (i < j && (uncapped || iter++ < max_iter)) {
		if (!uncapped && i < (j >> 1)) i = (j >> 1);       /* K/2 limit band size */

		const uint32_t start = limits[i];
		const uint32_t end = limits[j];
		const uint32_t ncond = limits[i + 1] - start;
		const REAL *od = pd[d];
		REAL *nd = pd[!d];

		#pragma omp parallel for schedule(static) num_threads(nthreads)
		for (int64_t vi = start; vi < (int64_t) end; vi++) {
			const uint32_t v = order ? order[vi] : (uint32_t) vi;
			if (pc) nd[v] = SFX(relax_cl)(m, od, pc[d], v, &pc[!d][v]);
			else nd[v] = SFX(relax)(m, od, v);
		}

		uint32_t count = 0;
		#pragma omp parallel for schedule(static) num_threads(nthreads) reduction(+: count)
		for (int64_t vi = start; vi < (int64_t) (start + ncond); vi++) {
			const uint32_t v = order ? order[vi] : (uint32_t) vi;
			err[vi] = FABS(nd[v] - od[v]) / od[v];
			count += (double) err[vi] < tol;
		}

		if (trace && sweeps < cap) {
			or_sweep *r = trace + sweeps;
			r->iter = sweeps + 1; r->i = i; r->j = j; r->start = start; r->end = end;
			r->ncond = ncond; r->count = count; r->pad = 0;
		}
		sweeps++;

		if (ncond == count) i++;
		if (j < nl - 1) j++;
		d = !d;
	}
	*nsw = sweeps;
	return (variant & OR_LAST_WRITE) ? pd[d] : pd[!d];
}

/* Driver: plain (sorted-indirected) or coalesced (renumbered and rebuilt,
 * src/geodesics_ptp.cpp:15-38) mesh, buffer init with sources at 0
 * (:51-59, :127-135), schedule, and the scatter of the result back to the
 * input numbering (:111-113; src/cuda/geodesics_ptp_coalescence.cu:84-86). */
static int SFX(ptp)(const or_che *mesh, const uint32_t *src, uint32_t ns,
                    const uint32_t *limits, uint32_t nl, const uint32_t *sorted,
                    int variant, double tol, int nthreads,
                    REAL *dist, or_sweep *trace, uint32_t cap, uint32_t *nsw, uint32_t *clusters)
{
	if (nthreads <= 0) nthreads = omp_get_max_threads();
	const uint32_t nv = mesh->nv;
	*nsw = 0;
	if (ns == 0 || nl < 2) return OR_EINVAL;
	for (uint32_t q = 0; q < ns; q++) if (src[q] >= nv) return OR_EINVAL;

	or_che cm;
	uint32_t *inv = NULL;
	const or_che *use = mesh;
	int rc;
	if (variant & OR_COALESCE) {
		const uint32_t nr = limits[nl - 1];
		inv = (uint32_t *) malloc(sizeof(uint32_t) * nv);
		double *V = (double *) malloc(sizeof(double) * 3 * (size_t) (nr ? nr : 1));
		uint32_t *F = (uint32_t *) malloc(sizeof(uint32_t) * (mesh->nhe ? mesh->nhe : 1));
		if (!inv || !V || !F) { free(inv); free(V); free(F); return OR_ENOMEM; }
		for (uint32_t v = 0; v < nv; v++) inv[v] = OR_NIL;
		for (uint32_t q = 0; q < nr; q++) {
			const uint32_t v = sorted[q];
			V[3 * (size_t) q + 0] = mesh->gt[3 * (size_t) v + 0];
			V[3 * (size_t) q + 1] = mesh->gt[3 * (size_t) v + 1];
			V[3 * (size_t) q + 2] = mesh->gt[3 * (size_t) v + 2];
			inv[v] = q;
		}
		uint32_t nF = 0;
		for (uint32_t he = 0; he < mesh->nhe; he++)
			if (inv[mesh->vt[he]] != OR_NIL && inv[mesh->vt[or_prev(he)]] != OR_NIL &&
			    inv[mesh->vt[or_next(he)]] != OR_NIL)
				F[nF++] = inv[mesh->vt[he]];
		rc = or_che_build(V, nr, F, nF / 3, &cm);
		free(V); free(F);
		if (rc != OR_OK) { free(inv); return rc; }
		use = &cm;
	}

	const uint32_t n = use->nv;
	REAL *gt = NULL;
	const REAL *gtp;
	if (sizeof(REAL) == sizeof(double)) {
		gtp = (const REAL *) use->gt;
	} else {
		gt = (REAL *) malloc(sizeof(REAL) * 3 * (size_t) n);
		if (!gt) { if (inv) { free(inv); or_che_free(&cm); } return OR_ENOMEM; }
		for (size_t q = 0; q < 3 * (size_t) n; q++) gt[q] = (REAL) use->gt[q];
		gtp = gt;
	}
	SFX(mview) mv = { n, gtp, use->vt, use->ot, use->evt };

	REAL *pd0 = (REAL *) malloc(sizeof(REAL) * n);
	REAL *pd1 = (REAL *) malloc(sizeof(REAL) * n);
	REAL *err = (REAL *) malloc(sizeof(REAL) * n);
	if (!pd0 || !pd1 || !err) {
		free(pd0); free(pd1); free(err); free(gt);
		if (inv) { free(inv); or_che_free(&cm); }
		return OR_ENOMEM;
	}
	for (uint32_t v = 0; v < n; v++) pd0[v] = pd1[v] = (REAL) INFINITY;
	/* Cluster labels: source q gets q + 1 (src/cuda/geodesics_ptp.cu:188-189);
	 * every other vertex starts at NIL (the reference leaves it unset). */
	uint32_t *pc[2] = { NULL, NULL };
	if (clusters) {
		pc[0] = (uint32_t *) malloc(sizeof(uint32_t) * n);
		pc[1] = (uint32_t *) malloc(sizeof(uint32_t) * n);
		if (!pc[0] || !pc[1]) {
			free(pc[0]); free(pc[1]); free(pd0); free(pd1); free(err); free(gt);
			if (inv) { free(inv); or_che_free(&cm); }
			return OR_ENOMEM;
		}
		for (uint32_t v = 0; v < n; v++) pc[0][v] = pc[1][v] = OR_NIL;
	}
	for (uint32_t q = 0; q < ns; q++) {
		const uint32_t s = inv ? inv[src[q]] : src[q];
		pd0[s] = pd1[s] = 0;
		if (clusters) pc[0][s] = pc[1][s] = q + 1;
	}

	REAL *res = SFX(schedule)(&mv, limits, nl, inv ? NULL : sorted, variant, tol, nthreads,
	                          pd0, pd1, err, trace, cap, nsw, clusters ? pc : NULL);
	const uint32_t *cres = clusters ? (res == pd0 ? pc[0] : pc[1]) : NULL;

	if (inv) {
		for (uint32_t v = 0; v < nv; v++) dist[v] = inv[v] != OR_NIL ? res[inv[v]] : (REAL) INFINITY;
		if (clusters)
			for (uint32_t v = 0; v < nv; v++) clusters[v] = inv[v] != OR_NIL ? cres[inv[v]] : OR_NIL;
	} else {
		memcpy(dist, res, sizeof(REAL) * nv);
		if (clusters) memcpy(clusters, cres, sizeof(uint32_t) * nv);
	}

	free(pc[0]); free(pc[1]);
	free(pd0); free(pd1); free(err); free(gt);
	if (inv) { free(inv); or_che_free(&cm); }
	return OR_OK;
}

/* iter_error_run_ptp_coalescence_gpu (src/cuda/test_geodesics_ptp_coalescence.cu:214-266)
 * on an already coalesced mesh (ring-numbered, exact distances in that
 * numbering): the uncapped loop; every sweep appends (sweep, i, j, window
 * size) to band; after every sweep run once j has reached the last ring,
 * (sweep, compute_error of the buffer it wrote) goes to iter/err.  Pinned with
 * the schedule (the loop is the uncapped one of SFX(schedule)). */
static int SFX(iter_error)(const SFX(mview) *m, const uint32_t *src_c, uint32_t ns, const uint32_t *limits,
                           uint32_t nl, const double *exact_c, double tol,
                           uint32_t *iter, double *errv, uint32_t cap, uint32_t *n_err,
                           uint32_t *band, uint32_t cap_band, uint32_t *n_band)
{
	const uint32_t n = m->nv;
	REAL *pd[2] = { (REAL *) malloc(sizeof(REAL) * n), (REAL *) malloc(sizeof(REAL) * n) };
	if (!pd[0] || !pd[1]) { free(pd[0]); free(pd[1]); return OR_ENOMEM; }
	for (uint32_t v = 0; v < n; v++) pd[0][v] = pd[1][v] = (REAL) INFINITY;
	for (uint32_t q = 0; q < ns; q++) pd[0][src_c[q]] = pd[1][src_c[q]] = 0;
	uint32_t d = 0, i = 1, j = 2, it = 0, ne = 0, nb = 0;
	while (nl >= 3 && i < j) {
		it++;
		const uint32_t start = limits[i], end = limits[j], ncond = limits[i + 1] - start;
		const REAL *od = pd[d];
		REAL *nd = pd[!d];
		for (uint32_t v = start; v < end; v++) nd[v] = SFX(relax)(m, od, v);
		if (nb < cap_band) { band[4 * nb] = it; band[4 * nb + 1] = i; band[4 * nb + 2] = j; band[4 * nb + 3] = end - start; }
		nb++;
		if (j == nl - 1) {
			double e = 0;
			for (uint32_t v = 0; v < n; v++)
				if (exact_c[v] > 0) e += fabs((double) nd[v] - exact_c[v]) / exact_c[v];
			if (ne < cap) { iter[ne] = it; errv[ne] = e * 100 / (double) (n - ns); }
			ne++;
		}
		uint32_t count = 0;
		for (uint32_t v = start; v < start + ncond; v++) count += (double) (FABS(nd[v] - od[v]) / od[v]) < tol;
		if (count == ncond) i++;
		if (j < nl - 1) j++;
		d = !d;
	}
	*n_err = ne;
	*n_band = nb;
	free(pd[0]); free(pd[1]);
	return OR_OK;
}

