// Per-vertex relaxation on the ELL slot layout, shared by the sweep kernels
// (ptp_gpu.hip) and the batched solver (batch.hip).  A is any argument struct
// with the fields cid, cgeo, cgeo64, cx64, h, inv_h.
#pragma once

#include <type_traits>

#include "device.hpp"
#include "update_step.hpp"

namespace gp {

constexpr uint32_t MAXS = 10;  // register-resident slots; longer fans take the streaming loop

// FP64 update of triangle (k, k+1) of the lane at slot base (flagged
// triangles).  ST is the distance between a lane's consecutive slots: 64 in
// the chunked layout, 1 in the vertex-major one.
template <uint32_t ST = 64, typename A>
__device__ __forceinline__ float tri_flagged(const A &a, size_t at, float ta, float tb)
{
	const double4 gd = a.cgeo64[at];
	const double q1 = a.cgeo64[at + ST].x;
	return (float) update_stable<double>(gd.x, q1, gd.y, gd.z, gd.w, (double) ta, (double) tb, (double) a.h,
	                                     (double) a.inv_h);
}

__device__ __forceinline__ double tri_f64(const double4 &xa, const double4 &xb, double ta, double tb)
{
	const double X0[3] = { xa.x, xa.y, xa.z };
	const double X1[3] = { xb.x, xb.y, xb.z };
	return update_ref_f64_x(X0, X1, ta, tb);
}

// New value of the vertex at lane slot base (chunk with S slots).  FP32:
// flagged triangles are skipped in the slot loop and done afterwards in a
// loop that runs as many times as the most-flagged lane of the wave needs,
// so the FP64 code runs ~1-2 times per wave instead of once per slot.
//
// CL: also report in *bk the triangle (slot k of triangle (k, k+1)) that set
// the new value -- the first in fan order among those reaching the minimum,
// NO_TRI if the value did not drop -- for the cluster label.
constexpr uint32_t NO_TRI = 0xFFFFFFFFu;

template <bool F64, uint32_t ST = 64, bool PF = true, bool CL = false, typename A>
__device__ __forceinline__ typename std::conditional<F64, double, float>::type
relax_lane(const A &a, size_t base, uint32_t S, const typename std::conditional<F64, double, float>::type *od,
           typename std::conditional<F64, double, float>::type best, uint32_t *bk_out = nullptr)
{
	using T = typename std::conditional<F64, double, float>::type;
	using G = typename std::conditional<F64, double4, float4>::type;
	const T old = best;
	uint32_t bk = NO_TRI;
	// Strictly lower wins; an equal value from an earlier triangle (evaluated
	// later because FP64 triangles are deferred) takes the label, as the
	// in-order scan of the reference would have kept it.
	auto take = [&](T p, uint32_t k) {
		if (p < best) {
			best = p;
			if constexpr (CL) bk = k;
		} else if constexpr (CL) {
			if (p == best && p < old && k < bk) bk = k;
		}
	};
	const G *geo;
	if constexpr (F64) geo = a.cx64; else geo = a.cgeo;
	uint32_t fmask = 0;
	if (S <= MAXS) {
		uint32_t id[MAXS];
		G g[MAXS];
		T t[MAXS];
#pragma unroll
		for (uint32_t k = 0; k < MAXS; k++) {
			id[k] = SLOT_EMPTY;
			if (k < S) { id[k] = a.cid[base + ST * k]; g[k] = geo[base + ST * k]; }
		}
		// FP32 (PF): the FP64 geometry of the first flagged triangle is
		// requested with the neighbour distances (both depend only on the slot
		// ids), so it costs no extra memory round trip; later flagged
		// triangles of the lane (rare) are deferred to the loop below.  PF =
		// false (register-tight callers) defers them all.
		uint32_t k0 = MAXS;
		double4 gd0 = { 0, 0, 0, 0 };
		double q10 = 0;
		if constexpr (!F64 && PF) {
#pragma unroll
			for (uint32_t k = MAXS; k-- > 0;)
				if (k + 1 < S && (id[k] & F64_FLAG)) k0 = k;
			if (k0 != MAXS) {
				gd0 = a.cgeo64[base + ST * (size_t) k0];
				q10 = a.cgeo64[base + ST * (size_t) (k0 + 1)].x;
			}
		}
#pragma unroll
		for (uint32_t k = 0; k < MAXS; k++)
			t[k] = (k < S) ? od[id[k] & ID30_MASK] : (T) INFINITY;
		T fa = (T) INFINITY, fb = (T) INFINITY;
#pragma unroll
		for (uint32_t k = 0; k + 1 < MAXS; k++) {
			if constexpr (F64) {
				if (k + 1 < S && !(id[k + 1] & SLOT_EMPTY)) take(tri_f64(g[k], g[k + 1], t[k], t[k + 1]), k);
			} else if (k + 1 < S) {
				// Straight-line per triangle: every lane evaluates the FP32
				// update and keeps it only for a real, unflagged triangle;
				// flagged ones are routed to the FP64 update by selects.
				const bool real = !(id[k + 1] & SLOT_EMPTY);
				const bool fl = (id[k] & F64_FLAG) != 0u;
				const float p = update_stable<float>(g[k].x, g[k + 1].x, g[k].y, g[k].z, g[k].w, t[k], t[k + 1], a.h,
				                                     a.inv_h);
				take((real && !fl) ? p : (T) INFINITY, k);
				const bool first = real && fl && k == k0;
				fa = first ? t[k] : fa;
				fb = first ? t[k + 1] : fb;
				fmask |= (real && fl && k != k0) ? (1u << k) : 0u;
			}
		}
		if constexpr (!F64) {
			if (k0 != MAXS) {
				take((T) update_stable<double>(gd0.x, q10, gd0.y, gd0.z, gd0.w, (double) fa, (double) fb,
				                               (double) a.h, (double) a.inv_h), k0);
			}
		}
	} else {
		uint32_t ida = a.cid[base];
		G ga = geo[base];
		T ta = od[ida & ID30_MASK];
		for (uint32_t k = 1; k < S; k++) {
			const uint32_t idb = a.cid[base + ST * k];
			if (idb & SLOT_EMPTY) break;
			const G gb = geo[base + ST * k];
			const T tb = od[idb & ID30_MASK];
			if constexpr (F64) {
				take(tri_f64(ga, gb, ta, tb), k - 1);
			} else {
				const T p = (ida & F64_FLAG) ? tri_flagged<ST>(a, base + ST * (k - 1), ta, tb)
				                             : update_stable<float>(ga.x, gb.x, ga.y, ga.z, ga.w, ta, tb, a.h, a.inv_h);
				take(p, k - 1);
			}
			ida = idb; ga = gb; ta = tb;
		}
	}
	if constexpr (!F64) {
		while (fmask) {
			const uint32_t k = __builtin_ctz(fmask);
			fmask &= fmask - 1u;
			const size_t at = base + ST * (size_t) k;
			const float ta = od[a.cid[at] & ID30_MASK];
			const float tb = od[a.cid[at + ST] & ID30_MASK];
			take(tri_flagged<ST>(a, at, ta, tb), k);
		}
	}
	if constexpr (CL) *bk_out = bk;
	return best;
}

}  // namespace gp
