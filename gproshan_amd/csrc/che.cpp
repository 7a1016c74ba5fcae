// Compact half-edge mesh, triangle fans and host toplesets.
//
// The tables follow the reference's construction rules exactly
// (che::init -> update_evt_ot_et, src/che.cpp:1254-1264,1295-1362): VT is the
// face array; OT pairs a half-edge with the first unpaired reversed half-edge
// in the reference's scan order; EVT is the largest half-edge leaving v,
// replaced by the outgoing border half-edge on a border, NIL on a vertex with
// two border fans.  The construction here is parallel: a reversed pair that
// is the only candidate of both half-edges is order-independent and is paired
// directly; only the few half-edges with several candidates (non-manifold
// edges) are paired serially in the reference's order.
#include "gp_internal.hpp"

#include <omp.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <cstdarg>
#include <cstdio>
#include <cstring>

namespace gp {

static thread_local std::string g_last_error;

int fail(int status, const char *fmt, ...)
{
	char buf[512];
	va_list ap;
	va_start(ap, fmt);
	vsnprintf(buf, sizeof(buf), fmt, ap);
	va_end(ap);
	g_last_error = buf;
	return status;
}

const char *last_error() { return g_last_error.c_str(); }
void clear_error() { g_last_error.clear(); }

Che::~Che()
{
	if (dev) device_mesh_free(dev);
}

int build_che(const double *xyz, uint32_t nv, const uint32_t *faces, uint32_t nf, Che &m)
{
	if (nv == 0 || !xyz) return fail(GP_ERR_INVALID, "che: no vertices");
	if (nf && !faces) return fail(GP_ERR_INVALID, "che: faces is null");
	if (3ull * nf >= 0x7FFFFFFFull || nv >= 0x7FFFFFFFu)
		return fail(GP_ERR_CAPACITY, "che: more than 2^31 half-edges or vertices");
	const uint32_t nhe = 3u * nf;
	for (uint32_t he = 0; he < nhe; he++)
		if (faces[he] >= nv) return fail(GP_ERR_INVALID, "che: face index %u out of range", faces[he]);

	m.nv = nv; m.nf = nf; m.nhe = nhe; m.manifold = true;
	m.gt.assign(xyz, xyz + 3 * (size_t) nv);
	m.vt.assign(faces, faces + nhe);
	m.ot.assign(nhe, NIL);
	m.evt.assign(nv, NIL);

	// Half-edges leaving each vertex, ascending (counting sort).
	std::vector<uint32_t> off(nv + 1, 0), lst(nhe);
	for (uint32_t he = 0; he < nhe; he++) off[m.vt[he] + 1]++;
	for (uint32_t v = 0; v < nv; v++) off[v + 1] += off[v];
	{
		std::vector<uint32_t> fill(off.begin(), off.end() - 1);
		for (uint32_t he = 0; he < nhe; he++) lst[fill[m.vt[he]]++] = he;
	}

	// Candidates of he: g = prev(h), h leaving VT[he], with VT[g] == VT[next(he)].
	// The relation is symmetric (g is a candidate of he iff he is one of g).
	std::vector<uint8_t> ncand(nhe);
	std::vector<uint32_t> first(nhe, NIL);
	#pragma omp parallel for schedule(static)
	for (int64_t he = 0; he < (int64_t) nhe; he++) {
		const uint32_t a = m.vt[he], b = m.vt[he_next((uint32_t) he)];
		uint32_t c = 0, f = NIL;
		for (uint32_t q = off[a]; q < off[a + 1]; q++) {
			const uint32_t g = he_prev(lst[q]);
			if (m.vt[g] == b) { if (c == 0) f = g; c++; }
		}
		ncand[he] = (uint8_t) std::min<uint32_t>(c, 2);
		first[he] = f;
	}
	// Mutually unique pairs.
	#pragma omp parallel for schedule(static)
	for (int64_t he = 0; he < (int64_t) nhe; he++) {
		if (ncand[he] == 1) {
			const uint32_t g = first[he];
			if (ncand[g] == 1) m.ot[he] = g;
		}
	}
	// Several candidates: the reference's greedy scan, in ascending order.
	for (uint32_t he = 0; he < nhe; he++) {
		if (ncand[he] < 2 || m.ot[he] != NIL) continue;
		const uint32_t a = m.vt[he], b = m.vt[he_next(he)];
		for (uint32_t q = off[a]; q < off[a + 1]; q++) {
			const uint32_t g = he_prev(lst[q]);
			if (m.vt[g] == b && m.ot[he] == NIL && m.ot[g] == NIL) {
				m.ot[he] = g;
				m.ot[g] = he;
			}
		}
	}
	// ET holds he when it was still unpaired at its turn: unpaired, or paired
	// with a larger half-edge.
	uint32_t ne = 0;
	#pragma omp parallel for reduction(+: ne)
	for (int64_t he = 0; he < (int64_t) nhe; he++)
		ne += (m.ot[he] == NIL || m.ot[he] > (uint32_t) he);
	m.ne = ne;

	// EVT: largest outgoing half-edge, then the border rule per vertex.
	bool manifold = true;
	#pragma omp parallel for schedule(static) reduction(&&: manifold)
	for (int64_t v = 0; v < (int64_t) nv; v++) {
		if (off[v] == off[v + 1]) continue;
		uint32_t e = lst[off[v + 1] - 1];
		for (uint32_t q = off[v]; q < off[v + 1]; q++) {
			const uint32_t he = lst[q];
			if (m.ot[he] != NIL || e == NIL) continue;
			if (m.ot[e] == NIL && e != he) { manifold = false; e = NIL; }
			else e = he;
		}
		m.evt[v] = e;
	}
	m.manifold = manifold;

	// Fans.
	m.fan_off.assign(nv + 1, 0);
	m.fan_closed.assign(nv, 0);
	#pragma omp parallel for schedule(static)
	for (int64_t v = 0; v < (int64_t) nv; v++) {
		const uint32_t stop = m.evt[v];
		if (stop == NIL) continue;
		uint32_t c = 0, he = stop, guard = 0;
		bool closed = false;
		while (he != NIL && guard++ <= nhe) {
			c++;
			const uint32_t h2 = m.ot[he_prev(he)];
			if (h2 == stop) { closed = true; break; }
			he = h2;
		}
		m.fan_off[v + 1] = c + 1;
		m.fan_closed[v] = closed;
	}
	for (uint32_t v = 0; v < nv; v++) m.fan_off[v + 1] += m.fan_off[v];
	m.fan.assign(m.fan_off[nv], NIL);
	// Mean edge length: per-vertex sums added in vertex order, so the value
	// (it sets the FP32 unit h and the band widths) does not depend on the
	// thread count.
	std::vector<double> vsum(nv, 0.0);
	std::vector<uint32_t> vcnt(nv, 0);
	#pragma omp parallel for schedule(static)
	for (int64_t v = 0; v < (int64_t) nv; v++) {
		const uint32_t stop = m.evt[v];
		if (stop == NIL) continue;
		uint32_t *L = m.fan.data() + m.fan_off[v];
		const uint32_t c = m.fan_off[v + 1] - m.fan_off[v] - 1;
		uint32_t he = stop;
		for (uint32_t k = 0; k < c; k++) {
			L[k] = m.vt[he_next(he)];
			if (k + 1 == c) L[c] = m.vt[he_prev(he)];
			he = m.ot[he_prev(he)];
		}
		const double *p = &m.gt[3 * (size_t) v];
		for (uint32_t k = 0; k <= c; k++) {
			const double *q = &m.gt[3 * (size_t) L[k]];
			vsum[v] += std::sqrt((q[0] - p[0]) * (q[0] - p[0]) + (q[1] - p[1]) * (q[1] - p[1]) +
			                     (q[2] - p[2]) * (q[2] - p[2]));
			vcnt[v]++;
		}
	}
	double esum = 0;
	uint64_t ecnt = 0;
	for (uint32_t v = 0; v < nv; v++) { esum += vsum[v]; ecnt += vcnt[v]; }
	m.mean_edge = (ecnt && esum > 0) ? esum / (double) ecnt : 1.0;
	return GP_OK;
}

bool prefer_f64(const Che &m)
{
	// The counts are cached in the mesh on first use (several threads may ask).
	static std::mutex mu;
	std::lock_guard<std::mutex> lock(mu);
	if (m.n_flag_tri < 0) {
		long long nflag = 0, nslot = 0;
		#pragma omp parallel for schedule(static) reduction(+ : nflag, nslot)
		for (int64_t vv = 0; vv < (int64_t) m.nv; vv++) {
			const uint32_t v = (uint32_t) vv;
			const uint32_t s0 = m.fan_off[v], n = m.fan_off[v + 1] - s0;
			nslot += (n && m.fan_closed[v]) ? n - 1 : n;
			const double *pv = &m.gt[3 * (size_t) v];
			for (uint32_t k = 0; k + 1 < n; k++) {
				const double *pa = &m.gt[3 * (size_t) m.fan[s0 + k]];
				const double *pb = &m.gt[3 * (size_t) m.fan[s0 + k + 1]];
				const double X0[3] = { pa[0] - pv[0], pa[1] - pv[1], pa[2] - pv[2] };
				const double X1[3] = { pb[0] - pv[0], pb[1] - pv[1], pb[2] - pv[2] };
				const double q0 = X0[0] * X0[0] + X0[1] * X0[1] + X0[2] * X0[2];
				const double q1 = X1[0] * X1[0] + X1[1] * X1[1] + X1[2] * X1[2];
				const double g = X0[0] * X1[0] + X0[1] * X1[1] + X0[2] * X1[2];
				const double cx = X0[1] * X1[2] - X0[2] * X1[1];
				const double cy = X0[2] * X1[0] - X0[0] * X1[2];
				const double cz = X0[0] * X1[1] - X0[1] * X1[0];
				const double det = cx * cx + cy * cy + cz * cz;
				const double qq = q0 * q1;
				if (q0 * q1 - g * g == 0 || !(qq > 0) || det < FLAG_SIN2_MIN * qq || g < FLAG_COS_MIN * std::sqrt(qq) ||
				    q0 > FLAG_Q_RATIO * q1 || q1 > FLAG_Q_RATIO * q0)
					nflag++;
			}
		}
		m.n_fan_slots = nslot;
		m.n_flag_tri = nflag;
	}
	return 4 * m.n_flag_tri > m.n_fan_slots;
}

int build_layout(const Che &m, const uint32_t *order, uint32_t nr, bool f64, Layout &L, uint32_t tile, uint32_t W,
                 uint32_t uniform_max, bool wrap)
{
	const double ih = 1.0 / m.mean_edge, ih2 = ih * ih;
	if (W != 64 && W != 1) return fail(GP_ERR_INVALID, "layout: chunk width must be 64 or 1");
	// Slots of a vertex: its fan entries, less the repeated first entry of a
	// closed fan when wrapping.
	auto slots = [&](uint32_t v) {
		const uint32_t n = m.fan_off[v + 1] - m.fan_off[v];
		return (wrap && n && m.fan_closed[v]) ? n - 1 : n;
	};
	L.nchunks = (nr + W - 1) / W;
	L.coff.assign((size_t) L.nchunks + 1, 0);
	L.fcum.assign((size_t) nr + 1, 0);
	L.max_slots = 0;
	L.n_flagged = 0;
	for (uint32_t c = 0; c < L.nchunks; c++) {
		uint32_t S = 0;
		for (uint32_t pos = W * c; pos < std::min(nr, W * c + W); pos++) {
			const uint32_t v = order ? order[pos] : pos;
			if (v >= m.nv) return fail(GP_ERR_INVALID, "layout: vertex %u out of range", v);
			if (v >= SLOT_WRAP) return fail(GP_ERR_CAPACITY, "layout: vertex ids must stay below 2^29");
			S = std::max(S, slots(v));
		}
		L.coff[c + 1] = L.coff[c] + S;
		L.max_slots = std::max(L.max_slots, S);
	}
	L.uniform = 0;
	if (uniform_max && W == 64 && L.max_slots <= uniform_max) {
		for (uint32_t c = 0; c <= L.nchunks; c++) L.coff[c] = c * L.max_slots;
		L.uniform = L.max_slots;
	}
	const size_t lanes = (size_t) L.coff[L.nchunks] * W;
	L.cid.assign(std::max<size_t>(lanes, 1), 0);
	if (f64) {
		L.cx64.assign(std::max<size_t>(lanes, 1), Edge64{0, 0, 0});
		L.cgeo.clear(); L.cgeo64.clear();
	} else {
		L.cgeo.assign(std::max<size_t>(lanes, 1), Geo32{0.f, 0.f});
		L.cgeo64.assign(std::max<size_t>(lanes, 1), Geo64{0, 0, 0, 0});
		L.cx64.clear();
	}
	std::vector<uint32_t> flagged_at(nr, 0);
	#pragma omp parallel for schedule(dynamic, 64)
	for (int64_t pp = 0; pp < (int64_t) nr; pp++) {
		const uint32_t pos = (uint32_t) pp;
		const uint32_t c = pos / W, lane = pos % W;
		const uint32_t S = L.coff[c + 1] - L.coff[c];
		const size_t base = (size_t) L.coff[c] * W + lane;
		const uint32_t v = order ? order[pos] : pos;
		const uint32_t s0 = m.fan_off[v], n = m.fan_off[v + 1] - s0, ns = slots(v);
		const double *pv = &m.gt[3 * (size_t) v];
		uint32_t nflag = 0;
		for (uint32_t k = 0; k < S; k++) {
			const size_t at = base + W * (size_t) k;
			if (k >= ns) { L.cid[at] = v | SLOT_EMPTY; continue; }
			const uint32_t a = m.fan[s0 + k];
			const double *pa = &m.gt[3 * (size_t) a];
			// Same subtraction as the reference's vertex::operator-.
			const double X0[3] = { pa[0] - pv[0], pa[1] - pv[1], pa[2] - pv[2] };
			uint32_t id = a;
			if (f64) {
				L.cx64[at] = Edge64{ X0[0], X0[1], X0[2] };
			} else {
				const double q0 = X0[0] * X0[0] + X0[1] * X0[1] + X0[2] * X0[2];
				Geo64 gd{ q0 * ih2, 0, 0, 0 };
				if (k + 1 < n) {
					const double *pb = &m.gt[3 * (size_t) m.fan[s0 + k + 1]];
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
					// The reference's determinant q00 q11 - q01 q10 (its
					// operations, in its order) is exactly 0 for a zero-length
					// edge or exactly collinear edges: its update is NaN there
					// and the triangle never wins.  det = NaN in the table gives
					// the FP64 fallback the same outcome (the edge paths still
					// apply when an input is INF, as in the reference).
					const bool ref_degenerate = q0 * q1 - g * g == 0;
					if (ref_degenerate) gd.z = std::nan("");
					if (ref_degenerate || !(qq > 0) || det < FLAG_SIN2_MIN * qq || g < FLAG_COS_MIN * std::sqrt(qq) ||
					    q0 > FLAG_Q_RATIO * q1 || q1 > FLAG_Q_RATIO * q0) {
						id |= F64_FLAG;
						nflag++;
					}
				}
				L.cgeo[at] = Geo32{ (float) gd.x, (float) gd.y };
				L.cgeo64[at] = gd;
			}
			if (k == 0 && ns < n) id |= SLOT_WRAP;
			L.cid[at] = id;
		}
		flagged_at[pos] = nflag;
	}
	for (uint32_t pos = 0; pos < nr; pos++) L.fcum[pos + 1] = L.fcum[pos] + flagged_at[pos];
	L.n_flagged = L.fcum[nr];
	L.n_slots = 0;
	for (uint32_t pos = 0; pos < nr; pos++) L.n_slots += slots(order ? order[pos] : pos);
	L.lidx.clear();
	if (tile) {
		if (order) return fail(GP_ERR_INVALID, "layout: tiles need a mesh numbered by positions");
		L.lidx.assign(L.cid.size(), LOCAL_OUT);
		#pragma omp parallel for schedule(static)
		for (int64_t pp = 0; pp < (int64_t) nr; pp++) {
			const uint32_t pos = (uint32_t) pp, c = pos / W, lane = pos % W;
			const uint32_t S = L.coff[c + 1] - L.coff[c];
			for (uint32_t k = 0; k < S; k++) {
				const size_t at = (size_t) (L.coff[c] + k) * W + lane;
				const uint32_t id = L.cid[at];
				if (id & SLOT_EMPTY) continue;
				const uint32_t u = id & SLOT_ID_MASK;
				if (u / tile == pos / tile) L.lidx[at] = (uint16_t) (u % tile);
			}
		}
	}
	return GP_OK;
}

static inline uint32_t spread10(uint32_t x)
{
	x &= 0x3FF;
	x = (x | (x << 16)) & 0x030000FF;
	x = (x | (x << 8)) & 0x0300F00F;
	x = (x | (x << 4)) & 0x030C30C3;
	x = (x | (x << 2)) & 0x09249249;
	return x;
}

static inline uint64_t spread21(uint64_t x)
{
	x &= 0x1FFFFF;
	x = (x | (x << 32)) & 0x1F00000000FFFFull;
	x = (x | (x << 16)) & 0x1F0000FF0000FFull;
	x = (x | (x << 8)) & 0x100F00F00F00F00Full;
	x = (x | (x << 4)) & 0x10C30C30C30C30C3ull;
	x = (x | (x << 2)) & 0x1249249249249249ull;
	return x;
}

void morton_order(const Che &m, std::vector<uint32_t> &order)
{
	double lo[3] = { INFINITY, INFINITY, INFINITY }, hi[3] = { -INFINITY, -INFINITY, -INFINITY };
	for (uint32_t v = 0; v < m.nv; v++)
		for (int k = 0; k < 3; k++) {
			lo[k] = std::min(lo[k], m.gt[3 * (size_t) v + k]);
			hi[k] = std::max(hi[k], m.gt[3 * (size_t) v + k]);
		}
	const double ext = std::max({ hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2], 1e-300 });
	const double sc = 2097151.0 / ext;
	std::vector<std::pair<uint64_t, uint32_t>> keys(m.nv);
	#pragma omp parallel for schedule(static)
	for (int64_t vv = 0; vv < (int64_t) m.nv; vv++) {
		const uint32_t v = (uint32_t) vv;
		const double *p = &m.gt[3 * (size_t) v];
		uint64_t c = 0;
		for (int k = 0; k < 3; k++) c |= spread21((uint64_t) std::min(2097151.0, (p[k] - lo[k]) * sc)) << k;
		keys[v] = { c, v };
	}
	std::sort(keys.begin(), keys.end());
	order.resize(m.nv);
	for (uint32_t q = 0; q < m.nv; q++) order[q] = keys[q].second;
}

int band_order(const Che &m, const uint32_t *src, uint32_t ns, const uint32_t *levels, const double *dhat,
               double width, std::vector<uint32_t> &order, std::vector<uint32_t> &limits)
{
	if (!(width > 0)) return fail(GP_ERR_INVALID, "band width must be positive");
	double lo[3] = { INFINITY, INFINITY, INFINITY }, hi[3] = { -INFINITY, -INFINITY, -INFINITY };
	for (uint32_t v = 0; v < m.nv; v++)
		for (int k = 0; k < 3; k++) { lo[k] = std::min(lo[k], m.gt[3 * (size_t) v + k]); hi[k] = std::max(hi[k], m.gt[3 * (size_t) v + k]); }
	double sc[3];
	for (int k = 0; k < 3; k++) sc[k] = hi[k] > lo[k] ? 1023.0 / (hi[k] - lo[k]) : 0.0;
	std::vector<uint8_t> is_src(m.nv, 0);
	for (uint32_t q = 0; q < ns; q++) is_src[src[q]] = 1;
	// (band, morton, vertex) keys of the reached non-source vertices.
	std::vector<std::pair<uint64_t, uint32_t>> keys;
	keys.reserve(m.nv);
	for (uint32_t v = 0; v < m.nv; v++) {
		if (levels[v] == NIL || is_src[v]) continue;
		const double d = dhat[v];
		const double bq = std::isfinite(d) && d > 0 ? std::floor(d / width) : 0.0;
		const uint64_t band = 1 + (uint64_t) std::min(bq, 4.0e9);
		const double *p = &m.gt[3 * (size_t) v];
		const uint32_t mort = spread10((uint32_t) ((p[0] - lo[0]) * sc[0])) |
		                      (spread10((uint32_t) ((p[1] - lo[1]) * sc[1])) << 1) |
		                      (spread10((uint32_t) ((p[2] - lo[2]) * sc[2])) << 2);
		keys.emplace_back((band << 30) | mort, v);
	}
	std::sort(keys.begin(), keys.end());
	order.clear();
	order.reserve(ns + keys.size());
	for (uint32_t q = 0; q < ns; q++) order.push_back(src[q]);
	limits.assign(1, 0);
	uint64_t cur = ~0ull;
	for (size_t q = 0; q < keys.size(); q++) {
		const uint64_t band = keys[q].first >> 30;
		if (band != cur) { limits.push_back((uint32_t) order.size()); cur = band; }
		order.push_back(keys[q].second);
	}
	limits.push_back((uint32_t) order.size());
	return GP_OK;
}

int toplesets_host(const Che &m, const uint32_t *src, uint32_t ns, uint32_t *levels,
                   uint32_t *sorted, std::vector<uint32_t> &limits, uint32_t k)
{
	limits.clear();
	if (ns == 0) return GP_OK;                                   // src/che.cpp:548
	for (uint32_t q = 0; q < ns; q++)
		if (src[q] >= m.nv) return fail(GP_ERR_INVALID, "toplesets: source %u out of range", src[q]);
	std::fill(levels, levels + m.nv, NIL);
	uint32_t p = 0;
	for (uint32_t q = 0; q < ns; q++) {
		if (levels[src[q]] != NIL)
			return fail(GP_ERR_DUPLICATE_SOURCE, "toplesets: source %u listed twice", src[q]);
		sorted[p++] = src[q];
		levels[src[q]] = 0;
	}
	limits.push_back(0);
	uint32_t level = 0;
	for (uint32_t i = 0; i < p; i++) {
		const uint32_t v = sorted[i];
		if (levels[v] > level) {
			// With a level cut k the walk stops at the first vertex of level
			// k + 1, whose level-(k + 1) discoveries stay in sorted and in the
			// last ring (src/che.cpp:572-575, 592).
			level++;
			if (level > k) break;
			limits.push_back(i);
		}
		const uint32_t *L = m.fan.data() + m.fan_off[v];
		const uint32_t n = m.link_size(v);
		const uint32_t lv = levels[v] + 1;
		for (uint32_t k = 0; k < n; k++) {
			const uint32_t u = L[k];
			if (levels[u] == NIL) { levels[u] = lv; sorted[p++] = u; }
		}
	}
	limits.push_back(p);
	for (uint32_t i = p; i < m.nv; i++) sorted[i] = NIL;
	return GP_OK;
}

int coalesce(const Che &m, const uint32_t *sorted, uint32_t nr, Che &out, std::vector<uint32_t> &inv)
{
	inv.assign(m.nv, NIL);
	for (uint32_t q = 0; q < nr; q++) inv[sorted[q]] = q;
	// Faces with all three vertices reached, in their original order, with
	// renumbered ids (src/geodesics_ptp.cpp:33-35).
	std::vector<uint32_t> F;
	F.reserve(m.nhe);
	for (uint32_t f = 0; f < m.nf; f++) {
		const uint32_t a = inv[m.vt[3 * f]], b = inv[m.vt[3 * f + 1]], c = inv[m.vt[3 * f + 2]];
		if (a != NIL && b != NIL && c != NIL) { F.push_back(a); F.push_back(b); F.push_back(c); }
	}
	std::vector<double> V(3 * (size_t) nr);
	for (uint32_t q = 0; q < nr; q++)
		for (int k = 0; k < 3; k++) V[3 * (size_t) q + k] = m.gt[3 * (size_t) sorted[q] + k];
	if (nr == 0) return fail(GP_ERR_INVALID, "coalesce: nothing reached");
	return build_che(V.data(), nr, F.data(), (uint32_t) (F.size() / 3), out);
}

}  // namespace gp
