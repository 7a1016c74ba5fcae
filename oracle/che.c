/*
 * TEST INFRASTRUCTURE -- NOT PART OF THE PRODUCT.
 * Oracle: compact half-edge mesh construction, star/link and toplesets.
 *
 * Restates, step by step and in the reference's own iteration order:
 *   che::init(vertices, n_v, faces, n_f)  src/che.cpp:1254-1264
 *   che::update_evt_ot_et                 src/che.cpp:1295-1362
 *   trig / next / prev                    src/che.cpp:18-34
 *   for_star                              include/che.h:10
 *   che::link                             src/che.cpp:102-112
 *   che::compute_toplesets                src/che.cpp:546-593
 *
 * Pins (tests/test_oracle_che.py): hand-derived VT/OT/EVT tables for a
 * tetrahedron, an open 2x2 grid and a bowtie (two triangles sharing one
 * vertex), the CHE invariants OT[OT[h]] == h, VT[OT[h]] == VT[next(h)],
 * VT[next(OT[h])] == VT[h]; the EVT rule (largest outgoing half-edge of an
 * interior vertex, the outgoing border half-edge of a border vertex, NIL for
 * a vertex with two border fans); levels == graph hop distance (scipy BFS);
 * sorted == discovery order of an independent brute-force BFS written in
 * Python in the test itself.
 */
#include "oracle.h"

#include <stdlib.h>
#include <string.h>

uint32_t or_next(uint32_t he)
{
	if (he == OR_NIL) return OR_NIL;
	return 3u * (he / 3u) + (he + 1u) % 3u;
}

uint32_t or_prev(uint32_t he)
{
	if (he == OR_NIL) return OR_NIL;
	return 3u * (he / 3u) + (he + 2u) % 3u;
}

void or_che_free(or_che *m)
{
	if (!m) return;
	free(m->gt); free(m->vt); free(m->ot); free(m->evt); free(m->et);
	memset(m, 0, sizeof(*m));
}

int or_che_build(const double *xyz, uint32_t nv, const uint32_t *faces, uint32_t nf, or_che *m)
{
	memset(m, 0, sizeof(*m));
	if (nv == 0) return OR_EINVAL;
	const uint64_t nhe64 = 3ull * nf;
	if (nhe64 >= OR_NIL) return OR_EINVAL;
	const uint32_t nhe = (uint32_t) nhe64;
	for (uint32_t he = 0; he < nhe; he++)
		if (faces[he] >= nv) return OR_EINVAL;

	m->nv = nv; m->nf = nf; m->nhe = nhe; m->manifold = 1;
	m->gt = (double *) malloc(sizeof(double) * 3 * (size_t) nv);
	m->vt = (uint32_t *) malloc(sizeof(uint32_t) * (nhe ? nhe : 1));
	m->ot = (uint32_t *) malloc(sizeof(uint32_t) * (nhe ? nhe : 1));
	m->evt = (uint32_t *) malloc(sizeof(uint32_t) * nv);
	if (!m->gt || !m->vt || !m->ot || !m->evt) { or_che_free(m); return OR_ENOMEM; }
	memcpy(m->gt, xyz, sizeof(double) * 3 * (size_t) nv);
	if (nhe) memcpy(m->vt, faces, sizeof(uint32_t) * nhe);

	/* EVT starts NIL (src/che.cpp:1297) */
	for (uint32_t v = 0; v < nv; v++) m->evt[v] = OR_NIL;
	if (nf == 0) return OR_OK;

	/* EVT[v] = last half-edge leaving v; per-vertex list of half-edges in
	 * ascending order (src/che.cpp:1301-1308).  The list is kept as CSR. */
	uint32_t *cnt = (uint32_t *) calloc((size_t) nv + 1, sizeof(uint32_t));
	uint32_t *lst = (uint32_t *) malloc(sizeof(uint32_t) * nhe);
	uint32_t *et  = (uint32_t *) malloc(sizeof(uint32_t) * nhe);
	if (!cnt || !lst || !et) { free(cnt); free(lst); free(et); or_che_free(m); return OR_ENOMEM; }
	for (uint32_t he = 0; he < nhe; he++) {
		m->evt[m->vt[he]] = he;
		cnt[m->vt[he] + 1]++;
	}
	for (uint32_t v = 0; v < nv; v++) cnt[v + 1] += cnt[v];
	{
		uint32_t *fill = (uint32_t *) malloc(sizeof(uint32_t) * nv);
		if (!fill) { free(cnt); free(lst); free(et); or_che_free(m); return OR_ENOMEM; }
		memcpy(fill, cnt, sizeof(uint32_t) * nv);
		for (uint32_t he = 0; he < nhe; he++) lst[fill[m->vt[he]]++] = he;
		free(fill);
	}

	/* OT and ET: a half-edge he without an opposite opens an edge; it is
	 * paired with prev(h) for the first h leaving VT[he] (ascending) whose
	 * prev(h) starts at VT[next(he)] and is itself unpaired
	 * (src/che.cpp:1311-1331). */
	for (uint32_t he = 0; he < nhe; he++) m->ot[he] = OR_NIL;
	uint32_t ne = 0;
	for (uint32_t he = 0; he < nhe; he++) {
		if (m->ot[he] != OR_NIL) continue;
		et[ne++] = he;
		const uint32_t a = m->vt[he];
		const uint32_t b = m->vt[or_next(he)];
		for (uint32_t q = cnt[a]; q < cnt[a + 1]; q++) {
			const uint32_t h = lst[q];
			const uint32_t ph = or_prev(h);
			if (m->vt[ph] == b && m->ot[he] == OR_NIL && m->ot[ph] == OR_NIL) {
				m->ot[he] = ph;
				m->ot[ph] = he;
			}
		}
	}
	m->ne = ne;
	m->et = (uint32_t *) malloc(sizeof(uint32_t) * (ne ? ne : 1));
	if (m->et) memcpy(m->et, et, sizeof(uint32_t) * ne);

	/* Border rule: a border half-edge becomes EVT of its vertex unless the
	 * vertex already holds a different border half-edge, in which case the
	 * vertex is non-manifold and EVT becomes NIL (src/che.cpp:1343-1352). */
	for (uint32_t he = 0; he < nhe; he++) {
		const uint32_t v = m->vt[he];
		if (m->ot[he] == OR_NIL && m->evt[v] != OR_NIL) {
			if (m->ot[m->evt[v]] == OR_NIL && m->evt[v] != he) {
				m->manifold = 0;
				m->evt[v] = OR_NIL;
			} else {
				m->evt[v] = he;
			}
		}
	}

	free(cnt); free(lst); free(et);
	return m->et ? OR_OK : OR_ENOMEM;
}

/* for_star: he = EVT[v], then he <- OT[prev(he)] until NIL or back at the
 * start (include/che.h:10). */
uint32_t or_star(const or_che *m, uint32_t v, uint32_t *out, uint32_t cap)
{
	if (v >= m->nv) return 0;
	const uint32_t stop = m->evt[v];
	uint32_t n = 0;
	for (uint32_t he = stop; he != OR_NIL; ) {
		if (n < cap) out[n] = he;
		n++;
		const uint32_t h2 = m->ot[or_prev(he)];
		he = (h2 != stop) ? h2 : OR_NIL;
	}
	return n;
}

/* che::link: next(he) for every star half-edge, plus prev(he) when the
 * half-edge before it in the fan has no opposite (src/che.cpp:102-112).
 * Writes half-edge ids; the neighbour is VT of each. */
uint32_t or_link(const or_che *m, uint32_t v, uint32_t *out, uint32_t cap)
{
	if (v >= m->nv) return 0;
	const uint32_t stop = m->evt[v];
	uint32_t n = 0;
	for (uint32_t he = stop; he != OR_NIL; ) {
		if (n < cap) out[n] = or_next(he);
		n++;
		if (m->ot[or_prev(he)] == OR_NIL) {
			if (n < cap) out[n] = or_prev(he);
			n++;
		}
		const uint32_t h2 = m->ot[or_prev(he)];
		he = (h2 != stop) ? h2 : OR_NIL;
	}
	return n;
}

/* che::compute_toplesets (src/che.cpp:546-593). */
int or_toplesets(const or_che *m, const uint32_t *src, uint32_t ns, uint32_t k,
                 uint32_t *levels, uint32_t *sorted, uint32_t *limits,
                 uint32_t cap_limits, uint32_t *nlimits)
{
	*nlimits = 0;
	if (ns == 0) return OR_OK;                       /* src/che.cpp:548 */
	for (uint32_t q = 0; q < ns; q++) if (src[q] >= m->nv) return OR_EINVAL;

	for (uint32_t v = 0; v < m->nv; v++) levels[v] = OR_NIL;   /* :550 */

	uint32_t level = 0, p = 0, nl = 0;
	for (uint32_t q = 0; q < ns; q++) {                /* :554-561 */
		if (levels[src[q]] != OR_NIL) return OR_EDUPSOURCE;
		sorted[p++] = src[q];
		levels[src[q]] = level;
	}

	if (nl >= cap_limits) return OR_ECAPACITY;
	limits[nl++] = 0;                                  /* :563 */

	uint32_t lk[64];
	for (uint32_t i = 0; i < p; i++) {                 /* :564-589 */
		const uint32_t v = sorted[i];
		if (levels[v] > level) {
			level++;
			if (level > k) break;
			if (nl >= cap_limits) return OR_ECAPACITY;
			limits[nl++] = i;
		}
		const uint32_t n = or_link(m, v, lk, 64);
		if (n > 64) return OR_ECAPACITY;
		for (uint32_t q = 0; q < n; q++) {
			const uint32_t u = m->vt[lk[q]];
			if (levels[u] == OR_NIL) {
				levels[u] = levels[v] + 1;
				sorted[p++] = u;
			}
		}
	}
	if (nl >= cap_limits) return OR_ECAPACITY;
	limits[nl++] = p;                                  /* :592 */
	for (uint32_t i = p; i < m->nv; i++) sorted[i] = OR_NIL;  /* unreached: the front end's memset, src/geodesics.cpp:27 */
	*nlimits = nl;
	return OR_OK;
}
