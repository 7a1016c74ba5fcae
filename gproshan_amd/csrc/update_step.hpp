// Triangle update of the parallel toplesets propagation, shared by the host
// solver (ptp_cpu.cpp, compiled by g++ with -ffp-contract=off) and the gfx950
// kernels (compiled by hipcc; the FP64 form is evaluated with contraction off
// so that it rounds like the reference's default x86-64 build).
#pragma once

#include <cmath>
#include <cstdint>

#ifdef __HIPCC__
#define GP_HD __host__ __device__ __forceinline__
#else
#define GP_HD inline
#endif

namespace gp {

// The reference's planar-wavefront update of vertex v from the triangle
// (v, a, b) with arrival times t0 at a and t1 at b, given the edge vectors
// X0 = a - v and X1 = b - v (update_step, src/geodesics_ptp.cpp:201-262;
// cu_update_step, src/cuda/geodesics_ptp.cu:301-369).  Same operations in the
// same order; NaN (det == 0) is returned as NaN and loses every '<' in the
// caller, as in the reference.
//
// |X|^2 as the reference forms q00 and q11 (and the square under the edge
// paths' sqrt): the same expression, so a caller holding a whole fan can form
// it, and its square root, once per fan entry and pass them in
// (update_ref_f64_q) with the same bits as a result.
GP_HD double ref_sq(const double X[3])
{
#ifdef __clang__
#pragma clang fp contract(off)
#endif
	return X[0] * X[0] + X[1] * X[1] + X[2] * X[2];
}

template <bool PRE>
GP_HD double update_ref_core(const double X0[3], const double X1[3], double q00, double q11, double s0, double s1,
                             double t0, double t1)
{
#ifdef __clang__
#pragma clang fp contract(off)
#endif
	const double q01 = X0[0] * X1[0] + X0[1] * X1[1] + X0[2] * X1[2];
	const double q10 = X1[0] * X0[0] + X1[1] * X0[1] + X1[2] * X0[2];

	const double det = q00 * q11 - q01 * q10;
	const double Q00 = q11 / det;
	const double Q01 = -q01 / det;
	const double Q10 = -q10 / det;
	const double Q11 = q00 / det;

	const double s = Q00 + Q01 + Q10 + Q11;
	const double delta = t0 * (Q00 + Q10) + t1 * (Q01 + Q11);
	const double dis = delta * delta - s * (t0 * t0 * Q00 + t0 * t1 * (Q10 + Q01) + t1 * t1 * Q11 - 1);

	double p = (delta + sqrt(dis)) / s;

	const double tp0 = t0 - p, tp1 = t1 - p;
	const double n0 = tp0 * (X0[0] * Q00 + X1[0] * Q10) + tp1 * (X0[0] * Q01 + X1[0] * Q11);
	const double n1 = tp0 * (X0[1] * Q00 + X1[1] * Q10) + tp1 * (X0[1] * Q01 + X1[1] * Q11);
	const double n2 = tp0 * (X0[2] * Q00 + X1[2] * Q10) + tp1 * (X0[2] * Q01 + X1[2] * Q11);
	const double cond0 = X0[0] * n0 + X0[1] * n1 + X0[2] * n2;
	const double cond1 = X1[0] * n0 + X1[1] * n1 + X1[2] * n2;
	const double c0 = cond0 * Q00 + cond1 * Q01;
	const double c1 = cond0 * Q10 + cond1 * Q11;

	if (t0 == INFINITY || t1 == INFINITY || dis < 0 || c0 >= 0 || c1 >= 0) {
		const double dp0 = t0 + (PRE ? s0 : sqrt(ref_sq(X0)));
		const double dp1 = t1 + (PRE ? s1 : sqrt(ref_sq(X1)));
		p = (dp1 < dp0) ? dp1 : dp0;
	}
	return p;
}

GP_HD double update_ref_f64_x(const double X0[3], const double X1[3], double t0, double t1)
{
	return update_ref_core<false>(X0, X1, ref_sq(X0), ref_sq(X1), 0.0, 0.0, t0, t1);
}

// The same with q00 = ref_sq(X0), q11 = ref_sq(X1), s0 = sqrt(q00) and
// s1 = sqrt(q11) supplied by the caller.
GP_HD double update_ref_f64_q(const double X0[3], const double X1[3], double q00, double q11, double s0, double s1,
                              double t0, double t1)
{
	return update_ref_core<true>(X0, X1, q00, q11, s0, s1, t0, t1);
}

// The same from the three positions: X0 = ga - gv, X1 = gb - gv, as the
// reference subtracts them (vertex::operator-, src/vertex.cpp).
GP_HD double update_ref_f64(const double *gv, const double *ga, const double *gb, double t0, double t1)
{
#ifdef __clang__
#pragma clang fp contract(off)
#endif
	const double X0[3] = { ga[0] - gv[0], ga[1] - gv[1], ga[2] - gv[2] };
	const double X1[3] = { gb[0] - gv[0], gb[1] - gv[1], gb[2] - gv[2] };
	return update_ref_f64_x(X0, X1, t0, t1);
}

// Per-triangle geometry, precomputed on the host in FP64 in units of the
// mesh's mean edge length h:  q0 = |X0|^2, q1 = |X1|^2 (from the next fan
// entry), g = X0.X1, det = |X0 x X1|^2 (cross product, not q0 q1 - g^2),
// S = |X1 - X0|^2.
//
// The same update as update_ref_f64, rewritten so that FP32 keeps its
// precision: the quadratic (t - p1)^T Q (t - p1) = 1 is invariant under
// t -> t - m1, p -> p - m, so it is solved for u = (t - min t)/h, which is
// O(1), instead of t ~ distance/h^2 in Q-units where the discriminant
// cancels catastrophically.  Multiplying by det > 0 gives the adjugate form:
//   D = u0 (q1 - g) + u1 (q0 - g),  E = u0^2 q1 - 2 u0 u1 g + u1^2 q0 - det,
//   p' = (D + sqrt(D^2 - S E)) / S,
//   sign c0 = sign(q1 (u0 - p') - g (u1 - p')),  sign c1 = sign(q0 (u1 - p') - g (u0 - p')),
// which are the reference's delta, dis, p and c scaled by det (and det^2).
// T = float is the bulk path; T = double is the fallback for triangles whose
// angle at the updated vertex is near-degenerate or obtuse.
// Square root and reciprocal of the stable update.  The FP32 bulk path uses
// the hardware v_sqrt_f32 / v_rcp_f32 (1 ulp) on the GPU.
GP_HD float stable_sqrt(float x)
{
#ifdef __HIP_DEVICE_COMPILE__
	return __builtin_amdgcn_sqrtf(x);
#else
	return std::sqrt(x);
#endif
}
GP_HD double stable_sqrt(double x) { return sqrt(x); }
GP_HD float stable_rcp(float x)
{
#ifdef __HIP_DEVICE_COMPILE__
	return __builtin_amdgcn_rcpf(x);
#else
	return 1.0f / x;
#endif
}
GP_HD double stable_rcp(double x) { return 1.0 / x; }
GP_HD float stable_fma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
GP_HD double stable_fma(double a, double b, double c) { return __builtin_fma(a, b, c); }

// Branch-free: both the planar solution and the edge paths are evaluated and
// the result is selected, so the unrolled fan loops of the kernels run as
// straight-line code whose triangles the compiler can interleave.  The
// selection is the one of the branching form: the edge paths when an input is
// infinite, the discriminant is negative or the front does not arrive from
// inside the wedge (the planar terms computed from an infinite input are
// discarded by the same test).
//
// s0 = sqrt(q0) and s1 = sqrt(q1) (edge lengths in units of h) are taken as
// arguments, so that a caller holding a whole fan computes each slot's once.
//
// Contraction is off and the fused multiply-adds are written out, so every
// inlined copy rounds alike (the FP32-table and FP64-table register paths must
// give the same bits).
template <typename T>
GP_HD T update_stable_s(T s0, T s1, T q0, T q1, T g, T det, T S, T t0, T t1, T h, T inv_h)
{
#ifdef __clang__
#pragma clang fp contract(off)
#endif
	const T inf = (T) INFINITY;
	const T dp0 = stable_fma(s0, h, t0);
	const T dp1 = stable_fma(s1, h, t1);
	// fmin: one min instruction; the inputs are never NaN (a finite value or
	// INF plus a finite length) nor -0, where it could differ from a compare.
	const T edge = fmin(dp0, dp1);
	const T m = fmin(t0, t1);
	// One of u0, u1 is 0 (the earlier arrival): with du the other one, qa the
	// earlier vertex's q and qb the later one's, D = du (qa - g),
	// E = du^2 qa - det, and the two conditions become
	// qa du - p' (qa - g) (the later vertex's) and p' (g - qb) - g du (the
	// earlier one's) -- the terms above with the zero products dropped.
	const bool first = t0 <= t1;
	const T du = fabs(t1 - t0) * inv_h;
	const T qa = first ? q0 : q1;
	const T qb = first ? q1 : q0;
	const T ag = qa - g;
	const T D = du * ag;
	const T E = stable_fma(du * du, qa, -det);
	const T disc = stable_fma(D, D, -(S * E));
	const T pp = (D + stable_sqrt(disc)) * stable_rcp(S);
	const T cl = stable_fma(qa, du, -(pp * ag));
	const T ce = stable_fma(pp, g - qb, -(g * du));
	const bool planar = (t0 != inf && t1 != inf) && !(disc < (T) 0 || cl >= (T) 0 || ce >= (T) 0);
	return planar ? stable_fma(pp, h, m) : edge;
}

template <typename T>
GP_HD T update_stable(T q0, T q1, T g, T det, T S, T t0, T t1, T h, T inv_h)
{
	return update_stable_s<T>(stable_sqrt(q0), stable_sqrt(q1), q0, q1, g, det, S, t0, t1, h, inv_h);
}

// The FP32 bulk path from the compact slot table (q0 and g per slot, q1 from
// the next slot): det = q0 q1 - g^2 and S = q0 + q1 - 2 g are formed here in
// FP32, one fused multiply-add each, so a slot carries 8 bytes of geometry
// instead of 16.  Their rounding error relative to the exact values is
// about 2.4e-7 / sin^2(theta) for det and 2.4e-7 / (4 sin^2(theta / 2)) for S
// (theta: the angle at the updated vertex, q0 ~ q1); the corners where that
// grows (sin^2 < 0.01) are flagged and take the FP64 update on the FP64
// table, whose det and S come from the cross product and the third edge.
GP_HD float update_stable_qs(float s0, float s1, float q0, float q1, float g, float t0, float t1, float h, float inv_h)
{
#ifdef __clang__
#pragma clang fp contract(off)
#endif
	const float det = __builtin_fmaf(q0, q1, -(g * g));
	const float S = __builtin_fmaf(-2.f, g, q0 + q1);
	return update_stable_s<float>(s0, s1, q0, q1, g, det, S, t0, t1, h, inv_h);
}

GP_HD float update_stable_q(float q0, float q1, float g, float t0, float t1, float h, float inv_h)
{
	return update_stable_qs(stable_sqrt(q0), stable_sqrt(q1), q0, q1, g, t0, t1, h, inv_h);
}

}  // namespace gp
