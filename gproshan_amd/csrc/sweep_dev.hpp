// Per-vertex relaxation on the ELL slot layout, shared by the sweep kernels
// (ptp_gpu.hip) and the batched solver (batch.hip).  A is any argument struct
// with the fields cid, cgeo, cgeo64, cx64, h, inv_h.
#pragma once

#include <type_traits>

#include "device.hpp"
#include "update_step.hpp"

namespace gp {

constexpr uint32_t MAXS = 10;  // register-resident slots; longer fans take the streaming loop

// Raw buffer loads (the LLVM intrinsics, declared by name) through a
// descriptor over [p, p + bytes): a load at or past `bytes` returns 0 and
// touches no memory.  The descriptor's last word is the untyped-buffer format
// of gfx9-family parts (32-bit data format, stride 0, raw offsets).
typedef float f32x2_t __attribute__((ext_vector_type(2)));
typedef double f64x2_t __attribute__((ext_vector_type(2)));
__device__ int raw_load_i32(__amdgpu_buffer_rsrc_t r, int voff, int soff, int aux) __asm(
    "llvm.amdgcn.raw.ptr.buffer.load.i32");
__device__ f32x2_t raw_load_f32x2(__amdgpu_buffer_rsrc_t r, int voff, int soff, int aux) __asm(
    "llvm.amdgcn.raw.ptr.buffer.load.v2f32");
__device__ f64x2_t raw_load_f64x2(__amdgpu_buffer_rsrc_t r, int voff, int soff, int aux) __asm(
    "llvm.amdgcn.raw.ptr.buffer.load.v2f64");
__device__ double raw_load_f64(__amdgpu_buffer_rsrc_t r, int voff, int soff, int aux) __asm(
    "llvm.amdgcn.raw.ptr.buffer.load.f64");

template <typename T>
__device__ __forceinline__ __amdgpu_buffer_rsrc_t rows_rsrc(const T *p, uint32_t bytes)
{
	return __builtin_amdgcn_make_buffer_rsrc((void *) p, 0, (int) bytes, 0x00020000);
}

// FP64 update of a flagged triangle: slot at (its first vertex) and slot
// at_n (its second vertex: the next slot, or the lane's slot 0 for the
// triangle that closes a fan stored without its repeated first entry).
template <typename A>
__device__ __forceinline__ float tri_flagged(const A &a, size_t at, size_t at_n, float ta, float tb)
{
	const double4 gd = a.cgeo64[at];
	const double q1 = a.cgeo64[at_n].x;
	return (float) update_stable<double>(gd.x, q1, gd.y, gd.z, gd.w, (double) ta, (double) tb, (double) a.h,
	                                     (double) a.inv_h);
}

__device__ __forceinline__ double tri_f64(const Edge64 &xa, const Edge64 &xb, double ta, double tb)
{
	const double X0[3] = { xa.x, xa.y, xa.z };
	const double X1[3] = { xb.x, xb.y, xb.z };
	return update_ref_f64_x(X0, X1, ta, tb);
}

// The same with each fan entry's |X|^2 and |X| formed once by the caller.
__device__ __forceinline__ double tri_f64_q(const Edge64 &xa, const Edge64 &xb, double qa, double qb, double sa,
                                            double sb, double ta, double tb)
{
	const double X0[3] = { xa.x, xa.y, xa.z };
	const double X1[3] = { xb.x, xb.y, xb.z };
	return update_ref_f64_q(X0, X1, qa, qb, sa, sb, ta, tb);
}

__device__ __forceinline__ double edge_sq(const Edge64 &x)
{
	const double X[3] = { x.x, x.y, x.z };
	return ref_sq(X);
}

// New value of the vertex at lane slot base (chunk with S slots; fans of up to
// MS slots are held in registers, longer ones streamed; ST is the
// distance between a lane's consecutive slots: 64 in the chunked layout, 1 in
// the vertex-major one).  Triangle k joins slot k and the next slot, or slot 0
// after the last real slot when slot 0 carries SLOT_WRAP (a closed fan stored
// without its repeated first entry).  FP32: flagged triangles are skipped in
// the slot loop; the first one of each lane is evaluated in FP64 from
// geometry fetched with the neighbour distances, later ones (rare) in a loop
// that runs as many times as the most-flagged lane of the wave needs.
//
// CL: also report in *bk the triangle that set the new value -- the first in
// fan order among those reaching the minimum, NO_TRI if the value did not
// drop -- for the cluster label.
//
// G64 (FP32, register path): the slots' FP64 geometry is read instead of the
// FP32 one; an unflagged triangle runs the FP32 update on its FP32 rounding
// (the FP32 table holds exactly that rounding), a flagged one the FP64 update
// on the FP64 values (what the deferred path loads), so the value is the same
// bits with one 32-byte read per slot and no second pass.
constexpr uint32_t NO_TRI = 0xFFFFFFFFu;

// Distances of slots 0 .. S-1 of a lane (INF past S).  OD is a distance
// array here; the dataflow sweep passes an accessor with its own overload that
// waits until each neighbour's value of the previous sweep is published.
template <uint32_t MS, typename T>
__device__ __forceinline__ void gather_slots(const T *od, const uint32_t (&id)[MS], uint32_t S, T (&t)[MS])
{
#pragma unroll
	for (uint32_t k = 0; k < MS; k++) t[k] = (k < S) ? od[id[k] & SLOT_ID_MASK] : (T) INFINITY;
}

// FP32 relaxation of a lane whose fan is closed and stored without its
// repeat, with SC real slots and no flagged triangle: triangle k joins slots
// k and (k + 1) mod SC.  These are the values and the order of the general
// register path of relax_lane for such a lane, with each slot's edge length
// computed once and no per-slot selects.
template <uint32_t SC, uint32_t MS, typename A>
__device__ __forceinline__ float relax_closed(const A &a, const float2 (&g)[MS], const float (&t)[MS], float best)
{
	float s[SC];
#pragma unroll
	for (uint32_t k = 0; k < SC; k++) s[k] = stable_sqrt(g[k].x);
#pragma unroll
	for (uint32_t k = 0; k < SC; k++) {
		const uint32_t kn = k + 1 < SC ? k + 1 : 0;
		const float p = update_stable_qs(s[k], s[kn], g[k].x, g[kn].x, g[k].y, t[k], t[kn], a.h, a.inv_h);
		best = fminf(best, p);   // a NaN p (degenerate triangle) keeps best, as p < best would
	}
	return best;
}

template <bool F64, uint32_t ST = 64, bool PF = true, bool CL = false, uint32_t MS = MAXS, bool G64 = false,
          typename A, typename OD>
__device__ __forceinline__ typename std::conditional<F64, double, float>::type
relax_lane(const A &a, size_t base, uint32_t S, const OD &od, typename std::conditional<F64, double, float>::type best,
           uint32_t *bk_out = nullptr)
{
	using T = typename std::conditional<F64, double, float>::type;
	using G = typename std::conditional<F64, Edge64, float2>::type;
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
	if constexpr (!F64 && G64) {
		if (S <= MS) {
			uint32_t id[MS];
			double4 g[MS];
			T t[MS];
#pragma unroll
			for (uint32_t k = 0; k < MS; k++) {
				id[k] = SLOT_EMPTY;
				if (k < S) { id[k] = a.cid[base + ST * k]; g[k] = a.cgeo64[base + ST * k]; }
			}
			const bool wrap = (id[0] & SLOT_WRAP) != 0u;
			gather_slots(od, id, S, t);
#pragma unroll
			for (uint32_t k = 0; k < MS; k++) {
				if (k < S) {
					const bool nr = k + 1 < MS && k + 1 < S && !(id[(k + 1) % MS] & SLOT_EMPTY);
					const bool tri = !(id[k] & SLOT_EMPTY) && (nr || wrap);
					const bool fl = (id[k] & F64_FLAG) != 0u;
					const T tb = nr ? t[(k + 1) % MS] : t[0];
					const double q1 = nr ? g[(k + 1) % MS].x : g[0].x;
					T p = (T) INFINITY;
					if (tri && fl)
						p = (float) update_stable<double>(g[k].x, q1, g[k].y, g[k].z, g[k].w, (double) t[k], (double) tb,
						                                  (double) a.h, (double) a.inv_h);
					if (tri && !fl)
						p = update_stable_q((float) g[k].x, (float) q1, (float) g[k].y, t[k], tb, a.h, a.inv_h);
					take(p, k);
				}
			}
			if constexpr (CL) *bk_out = bk;
			return best;
		}
	}
	if (S <= MS) {
		// Slots past S hold id 0 (no flag bits); every use of id[k] below is
		// guarded by k < S.
		uint32_t id[MS];
		G g[MS];
		T t[MS];
		if constexpr (ST == 64) {
			// Chunked tables: all MS slot ids, then all MS geometry rows,
			// through descriptors that end at the chunk's S rows (a row past
			// S reads as 0, so id[k] = 0 there, with no branch and no memory
			// access).  The loads are straight-line and their count fixed, so
			// the closed-fan test and the neighbour gathers wait only for the
			// ids (loads return in order) while the geometry is arriving.
			// Every chunked caller passes base = (first slot row) * 64 + lane
			// and one S for the whole wave, so the descriptors are uniform.
			const uint32_t lane = threadIdx.x & 63u;
			const size_t r0 = base - lane;   // the chunk's first slot, the same in every lane
			const size_t rs = ((size_t) __builtin_amdgcn_readfirstlane((uint32_t) (r0 >> 32)) << 32) |
			                  __builtin_amdgcn_readfirstlane((uint32_t) r0);
			const __amdgpu_buffer_rsrc_t rid = rows_rsrc(a.cid + rs, S * 64u * 4u);
#pragma unroll
			for (uint32_t k = 0; k < MS; k++) id[k] = (uint32_t) raw_load_i32(rid, (int) (lane * 4u + k * 256u), 0, 0);
			const __amdgpu_buffer_rsrc_t rg = rows_rsrc(geo + rs, S * 64u * (uint32_t) sizeof(G));
#pragma unroll
			for (uint32_t k = 0; k < MS; k++) {
				const int o = (int) ((lane + 64u * k) * (uint32_t) sizeof(G));
				if constexpr (F64) {
					const f64x2_t xy = raw_load_f64x2(rg, o, 0, 0);
					g[k] = Edge64{ xy.x, xy.y, raw_load_f64(rg, o + 16, 0, 0) };
				} else {
					const f32x2_t v = raw_load_f32x2(rg, o, 0, 0);
					g[k] = make_float2(v.x, v.y);
				}
			}
		} else if constexpr (F64) {
			// Vertex-major tables (a lane's own S): every slot id before any
			// edge vector, slots past S rereading the lane's last row.
			if (S == 0) return best;
#pragma unroll
			for (uint32_t k = 0; k < MS; k++) id[k] = a.cid[base + ST * (k < S ? k : S - 1)];
#pragma unroll
			for (uint32_t k = 0; k < MS; k++) g[k] = geo[base + ST * (k < S ? k : S - 1)];
#pragma unroll
			for (uint32_t k = 0; k < MS; k++) id[k] = k < S ? id[k] : 0u;
		} else {
#pragma unroll
			for (uint32_t k = 0; k < MS; k++) {
				id[k] = 0;
				if (k < S) { id[k] = a.cid[base + ST * k]; g[k] = geo[base + ST * k]; }
			}
		}
		const bool wrap = (id[0] & SLOT_WRAP) != 0u;
		// FP32 without labels, chunked tables: when every active lane holds a
		// closed fan without empty slots or flagged triangles (the interior of
		// a mesh), the closed-fan path for the chunk's slot count.
		if constexpr (!F64 && !CL && ST == 64) {
			uint32_t flags = 0;
#pragma unroll
			for (uint32_t k = 0; k < MS; k++) flags |= id[k];
			const bool plain = wrap && !(flags & (SLOT_EMPTY | F64_FLAG));
			if (__ballot(!plain) == 0 && S >= 3 && S <= 8 && S <= MS) {
				gather_slots(od, id, S, t);
				switch (S) {
				case 3: return relax_closed<3>(a, g, t, best);
				case 4: return relax_closed<4>(a, g, t, best);
				case 5: return relax_closed<(MS >= 5 ? 5 : MS)>(a, g, t, best);
				case 6: return relax_closed<(MS >= 6 ? 6 : MS)>(a, g, t, best);
				case 7: return relax_closed<(MS >= 7 ? 7 : MS)>(a, g, t, best);
				default: return relax_closed<(MS >= 8 ? 8 : MS)>(a, g, t, best);
				}
			}
		}
		// Slot k + 1 is real (else the triangle of slot k closes onto slot 0).
		auto next_real = [&](uint32_t k) {
			return k + 1 < MS && k + 1 < S && !(id[(k + 1) % MS] & SLOT_EMPTY);
		};
		// FP32 (PF): the FP64 geometry of the first flagged triangle is
		// requested with the neighbour distances (both depend only on the slot
		// ids), so it costs no extra memory round trip; later flagged
		// triangles of the lane (rare) are deferred to the loop below.  PF =
		// false (register-tight callers) defers them all.
		uint32_t k0 = MS, kn0 = 0;
		double4 gd0 = { 0, 0, 0, 0 };
		double q10 = 0;
		if constexpr (!F64 && PF) {
#pragma unroll
			for (uint32_t k = MS; k-- > 0;)
				if (k < S && (id[k] & F64_FLAG)) { k0 = k; kn0 = next_real(k) ? k + 1 : 0; }
			if (k0 != MS) {
				gd0 = a.cgeo64[base + ST * (size_t) k0];
				q10 = a.cgeo64[base + ST * (size_t) kn0].x;
			}
		}
		gather_slots(od, id, S, t);
		T fa = (T) INFINITY, fb = (T) INFINITY;
		// FP64: each entry's |X|^2 and |X| once per fan (rolled from one
		// triangle to the next; entry 0's kept for the closing triangle).
		double q_0 = 0, s_0 = 0, q_a = 0, s_a = 0;
		if constexpr (F64) {
			q_0 = edge_sq(g[0]);
			s_0 = sqrt(q_0);
			q_a = q_0;
			s_a = s_0;
		}
#pragma unroll
		for (uint32_t k = 0; k < MS; k++) {
			if (k < S) {
				const bool nr = next_real(k);
				const bool tri = !(id[k] & SLOT_EMPTY) && (nr || wrap);
				const T tb = nr ? t[(k + 1) % MS] : t[0];
				if constexpr (F64) {
					const G gb = nr ? g[(k + 1) % MS] : g[0];
					double q_n = q_0, s_n = s_0;
					if (k + 1 < MS && k + 1 < S) {
						q_n = edge_sq(g[(k + 1) % MS]);
						s_n = sqrt(q_n);
					}
					const double q_b = nr ? q_n : q_0, s_b = nr ? s_n : s_0;
					if (tri) take(tri_f64_q(g[k], gb, q_a, q_b, s_a, s_b, t[k], tb), k);
					q_a = q_n;
					s_a = s_n;
				} else {
					// Straight-line per triangle: every lane evaluates the
					// FP32 update and keeps it only for a real, unflagged
					// triangle; flagged ones are routed to the FP64 update by
					// selects.
					const bool fl = (id[k] & F64_FLAG) != 0u;
					const float q1 = nr ? g[(k + 1) % MS].x : g[0].x;
					const float p = update_stable_q(g[k].x, q1, g[k].y, t[k], tb, a.h, a.inv_h);
					take((tri && !fl) ? p : (T) INFINITY, k);
					const bool first = tri && fl && k == k0;
					fa = first ? t[k] : fa;
					fb = first ? tb : fb;
					fmask |= (tri && fl && k != k0) ? (1u << k) : 0u;
				}
			}
		}
		if constexpr (!F64) {
			if (k0 != MS) {
				take((T) update_stable<double>(gd0.x, q10, gd0.y, gd0.z, gd0.w, (double) fa, (double) fb,
				                               (double) a.h, (double) a.inv_h), k0);
			}
		}
	} else {
		const uint32_t id0 = a.cid[base];
		const G g0 = geo[base];
		const T t0 = od[id0 & SLOT_ID_MASK];
		uint32_t ida = id0;
		G ga = g0;
		T ta = t0;
		uint32_t k = 1;
		auto tri = [&](const G &gb, T tb, size_t at_n, uint32_t kt) {
			if constexpr (F64) {
				take(tri_f64(ga, gb, ta, tb), kt);
			} else {
				const T p = (ida & F64_FLAG) ? tri_flagged(a, base + ST * (size_t) kt, at_n, ta, tb)
				                             : update_stable_q(ga.x, gb.x, ga.y, ta, tb, a.h, a.inv_h);
				take(p, kt);
			}
		};
		for (; k < S; k++) {
			const uint32_t idb = a.cid[base + ST * k];
			if (idb & SLOT_EMPTY) break;
			const G gb = geo[base + ST * k];
			const T tb = od[idb & SLOT_ID_MASK];
			tri(gb, tb, base + ST * (size_t) k, k - 1);
			ida = idb; ga = gb; ta = tb;
		}
		if (id0 & SLOT_WRAP) tri(g0, t0, base, k - 1);
	}
	if constexpr (!F64) {
		while (fmask) {
			const uint32_t k = __builtin_ctz(fmask);
			fmask &= fmask - 1u;
			const size_t at = base + ST * (size_t) k;
			const uint32_t idn = k + 1 < S ? a.cid[at + ST] : SLOT_EMPTY;
			const size_t at_n = (idn & SLOT_EMPTY) ? base : at + ST;
			const float ta = od[a.cid[at] & SLOT_ID_MASK];
			const float tb = od[a.cid[at_n] & SLOT_ID_MASK];
			take(tri_flagged(a, at, at_n, ta, tb), k);
		}
	}
	if constexpr (CL) *bk_out = bk;
	return best;
}

// Vertex-major slot tables of a mesh in its own numbering: slot k of vertex v
// at coff[v] + k (coff = the fan offsets), coff[v + 1] - coff[v] - 1 slots for
// a closed fan (stored without its repeated first entry), else
// coff[v + 1] - coff[v].
struct VertexTables {
	const uint32_t *coff;
	const uint8_t *closed;
	const uint32_t *cid;
	const float2 *cgeo;
	const double4 *cgeo64;
	const Edge64 *cx64;
};

// Vertex-major slot tables of a mesh in its own numbering (slot k of vertex v
// at fan_off[v] + k; a closed fan stored without its repeated first entry
// leaves its last fan entry unused), the source that device-built plans and
// farthest-point sampling gather their ring-ordered tables from.  Built on the
// device once per mesh and precision.
struct MasterLayout {
	uint32_t *cid = nullptr;
	float2 *cgeo = nullptr;
	double4 *cgeo64 = nullptr;
	Edge64 *cx64 = nullptr;
	uint32_t max_slots = 0;
	VertexTables view(const DeviceMesh *dm) const
	{
		return VertexTables{ dm->fan_off, dm->closed, cid, cgeo, cgeo64, cx64 };
	}
	~MasterLayout()
	{
		for (void *p : { (void *) cid, (void *) cgeo, (void *) cgeo64, (void *) cx64 }) (void) hipFree(p);
	}
};

// The mesh's vertex-major tables for the precision (built once per mesh) and
// its longest fan's slot count.
int master_tables(Che &mesh, bool f64, VertexTables &out, uint32_t &max_slots);

// FP64 geometry of a chunked slot read from the vertex-major tables through
// the slot's index there (flagged triangles are rare: no chunked copy).
struct GeoRemap64 {
	const double4 *geo;
	const uint32_t *midx;
	__device__ double4 operator[](size_t at) const { return geo[midx[at]]; }
};

// Ring-ordered chunked tables gathered from vertex-major ones: position p
// (vertex sorted[p], p < *nr_at) takes lane p % 64 of chunk p / 64, whose
// slot rows start at row coff[c] (c * S when coff is null); neighbour ids
// become positions (flags kept), missing slots are empty (id = p).  FP32
// tables keep the FP64 geometry as a copy (cgeo64) or, MIDX, as the slot's
// index in the vertex-major tables (midx).
template <bool F64, bool MIDX>
__global__ void k_gather_tables(const uint32_t *sorted, const uint32_t *inv, const uint32_t *nr_at, uint32_t npos,
                                uint32_t S, const uint32_t *coff, VertexTables m, uint32_t *cid, float2 *cgeo,
                                double4 *cgeo64, uint32_t *midx, Edge64 *cx64)
{
	const uint32_t nr = *nr_at;
	for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < npos; p += gridDim.x * blockDim.x) {
		const uint32_t c = p >> 6;
		const uint32_t row = coff ? coff[c] : c * S;
		const uint32_t Sc = coff ? coff[c + 1] - row : S;
		const size_t base = (size_t) row * 64 + (p & 63u);
		uint32_t mb = 0, n = 0;
		if (p < nr) {
			const uint32_t v = sorted[p];
			mb = m.coff[v];
			const uint32_t nf = m.coff[v + 1] - mb;
			n = (nf && m.closed[v]) ? nf - 1 : nf;
		}
		for (uint32_t k = 0; k < Sc; k++) {
			const size_t at = base + 64 * (size_t) k;
			const bool real = k < n;
			const uint32_t id = real ? m.cid[mb + k] : 0u;
			cid[at] = real ? (inv[id & SLOT_ID_MASK] | (id & ~SLOT_ID_MASK)) : (p | SLOT_EMPTY);
			if constexpr (F64) {
				cx64[at] = real ? m.cx64[mb + k] : Edge64{ 0, 0, 0 };
			} else {
				cgeo[at] = real ? m.cgeo[mb + k] : make_float2(0.f, 0.f);
				if constexpr (MIDX) midx[at] = real ? mb + k : 0u;
				else cgeo64[at] = real ? m.cgeo64[mb + k] : make_double4(0, 0, 0, 0);
			}
		}
	}
}

}  // namespace gp
