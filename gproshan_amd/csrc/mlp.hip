// Learned band predictor: per-vertex features, and a 10 -> 64 -> 64 -> 1 ReLU
// MLP with BF16 weights/activations and FP32 accumulation on gfx950 MFMA
// (v_mfma_f32_16x16x32_bf16).  No reference counterpart.
//
// MFMA operand layout (16x16x32, wave64; lane l, r = l & 15, g = l >> 4):
//   A (16 x 32): row r, k = 8g .. 8g+7   (8 bf16 per lane)
//   B (32 x 16): col r, k = 8g .. 8g+7
//   C (16 x 16): rows 4g .. 4g+3, col r (4 floats per lane)
// Every GEMM operand is read from LDS (or global) in a layout whose 8 k
// values are contiguous, so each fragment is one 16-byte load; the layer
// outputs are written back to LDS in the layout the next GEMM reads.
//
// Training is deterministic: each workgroup writes its gradient partial sums
// (no float atomics) and one kernel reduces them and applies Adam to the FP32
// master weights, then refreshes the BF16 copies.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <random>
#include <vector>

#include "device.hpp"

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

namespace gp {

constexpr uint32_t NF = 10;     // features
constexpr uint32_t KP = 32;     // padded input width
constexpr uint32_t HD = 64;     // hidden width
// FP32 master parameter vector: W1[64][32], b1[64], W2[64][64], b2[64], w3[64], b3[1].
constexpr uint32_t OFF_W1 = 0, OFF_B1 = OFF_W1 + HD * KP, OFF_W2 = OFF_B1 + HD, OFF_B2 = OFF_W2 + HD * HD,
                   OFF_W3 = OFF_B2 + HD, OFF_B3 = OFF_W3 + HD, NPARAM = OFF_B3 + 1;

struct MlpDev {
	float *w = nullptr;        // FP32 master parameters [NPARAM]
	float *m = nullptr, *v = nullptr;  // Adam moments
	__bf16 *w1 = nullptr;      // W1 [out 64][in 32]
	__bf16 *w2 = nullptr;      // W2 [out 64][in 64]
	__bf16 *w2t = nullptr;     // W2^T [in 64][out 64]
	uint32_t step = 0;
};

__device__ __forceinline__ floatx4 mfma(const bf16x8 &a, const bf16x8 &b, const floatx4 &c)
{
	return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ bf16x8 ld8(const __bf16 *p)
{
	return *(const bf16x8 *) p;
}

// Weights and biases staged in LDS by a workgroup.
struct WShared {
	__bf16 w1[HD * KP];
	__bf16 w2[HD * HD];
	__bf16 w2t[HD * HD];
	float b1[HD], b2[HD], w3[HD], b3;
};

__device__ void stage_weights(WShared &s, const float *w, const __bf16 *w1, const __bf16 *w2, const __bf16 *w2t,
                              bool need_t)
{
	for (uint32_t q = threadIdx.x; q < HD * KP; q += blockDim.x) s.w1[q] = w1[q];
	for (uint32_t q = threadIdx.x; q < HD * HD; q += blockDim.x) {
		s.w2[q] = w2[q];
		if (need_t) s.w2t[q] = w2t[q];
	}
	for (uint32_t q = threadIdx.x; q < HD; q += blockDim.x) {
		s.b1[q] = w[OFF_B1 + q];
		s.b2[q] = w[OFF_B2 + q];
		s.w3[q] = w[OFF_W3 + q];
	}
	if (threadIdx.x == 0) s.b3 = w[OFF_B3];
	__syncthreads();
}

// Forward of one 16-sample tile held by a wave.  xa: the A fragment of X.
// Writes h1 (bf16, [16][64] row-major) to hbuf; returns z2 + b2 (pre-ReLU)
// fragments in acc2 and y for rows 4g..4g+3 in yr (all 16 lanes of a group).
__device__ __forceinline__ void forward_tile(const WShared &s, const bf16x8 &xa, __bf16 *hbuf, floatx4 acc1[4],
                                             floatx4 acc2[4], float yr[4])
{
	const uint32_t l = threadIdx.x & 63u, r = l & 15u, g = l >> 4;
#pragma unroll
	for (uint32_t t = 0; t < 4; t++) {
		const floatx4 z = { 0.f, 0.f, 0.f, 0.f };
		acc1[t] = mfma(xa, ld8(&s.w1[(t * 16 + r) * KP + 8 * g]), z);
		const float bb = s.b1[t * 16 + r];
#pragma unroll
		for (uint32_t q = 0; q < 4; q++) {
			acc1[t][q] += bb;
			hbuf[(4 * g + q) * HD + t * 16 + r] = (__bf16) fmaxf(acc1[t][q], 0.f);
		}
	}
	__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
	__builtin_amdgcn_wave_barrier();
	__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
	const bf16x8 ha0 = ld8(&hbuf[r * HD + 8 * g]);
	const bf16x8 ha1 = ld8(&hbuf[r * HD + 32 + 8 * g]);
#pragma unroll
	for (uint32_t t = 0; t < 4; t++) {
		const floatx4 z = { 0.f, 0.f, 0.f, 0.f };
		floatx4 c = mfma(ha0, ld8(&s.w2[(t * 16 + r) * HD + 8 * g]), z);
		c = mfma(ha1, ld8(&s.w2[(t * 16 + r) * HD + 32 + 8 * g]), c);
		const float bb = s.b2[t * 16 + r];
#pragma unroll
		for (uint32_t q = 0; q < 4; q++) c[q] += bb;
		acc2[t] = c;
	}
#pragma unroll
	for (uint32_t q = 0; q < 4; q++) {
		float p = 0.f;
#pragma unroll
		for (uint32_t t = 0; t < 4; t++) p += fmaxf(acc2[t][q], 0.f) * s.w3[t * 16 + r];
		p += __shfl_xor(p, 8, 16);
		p += __shfl_xor(p, 4, 16);
		p += __shfl_xor(p, 2, 16);
		p += __shfl_xor(p, 1, 16);
		yr[q] = p + s.b3;
	}
}

// Inference: y[i] for rows of X ([n][32] bf16).
__global__ __launch_bounds__(256) void k_mlp_forward(const __bf16 *X, uint32_t n, const float *w, const __bf16 *w1,
                                                     const __bf16 *w2, float *y)
{
	__shared__ WShared s;
	__shared__ __bf16 hb[4][16 * HD];
	stage_weights(s, w, w1, w2, nullptr, false);
	const uint32_t l = threadIdx.x & 63u, r = l & 15u, g = l >> 4, wv = threadIdx.x >> 6;
	const uint32_t ntile = (n + 15) / 16;
	for (uint32_t tile = blockIdx.x * 4 + wv; tile < ntile; tile += gridDim.x * 4) {
		const uint32_t row = tile * 16 + r;
		bf16x8 xa;
		if (row < n) xa = ld8(&X[(size_t) row * KP + 8 * g]);
		else for (int q = 0; q < 8; q++) xa[q] = (__bf16) 0.f;
		floatx4 a1[4], a2[4];
		float yr[4];
		forward_tile(s, xa, hb[wv], a1, a2, yr);
		if (r == 0)
			for (uint32_t q = 0; q < 4; q++)
				if (tile * 16 + 4 * g + q < n) y[tile * 16 + 4 * g + q] = yr[q];
		__builtin_amdgcn_wave_barrier();
	}
}

// Gradient of mean((y - t)^2) over the samples idx[0..n) (rows of X), as
// per-workgroup partial sums partial[blockIdx][NPARAM]; loss partial sums in
// lpart[blockIdx].  scale = 2 / (number of samples of the whole batch).
__global__ __launch_bounds__(256) void k_mlp_grad(const __bf16 *X, const float *tgt, const uint32_t *idx, uint32_t n,
                                                  float scale, const float *w, const __bf16 *w1, const __bf16 *w2,
                                                  const __bf16 *w2t, float *partial, float *lpart)
{
	__shared__ WShared s;
	struct WaveBuf {
		__bf16 h1[32 * HD];    // [sample][in]
		__bf16 h1t[HD * 32];   // [in][sample]
		__bf16 dh2[32 * HD];   // [sample][out]
		__bf16 dh2t[HD * 32];  // [out][sample]
		__bf16 dh1t[HD * 32];  // [in][sample]
		__bf16 xt[KP * 32];    // [feature][sample]
	};
	__shared__ WaveBuf wb[4];
	__shared__ float red[NPARAM];
	__shared__ float lred;
	stage_weights(s, w, w1, w2, w2t, true);
	for (uint32_t q = threadIdx.x; q < NPARAM; q += blockDim.x) red[q] = 0.f;
	if (threadIdx.x == 0) lred = 0.f;
	__syncthreads();
	const uint32_t l = threadIdx.x & 63u, r = l & 15u, g = l >> 4, wv = threadIdx.x >> 6;
	WaveBuf &B = wb[wv];
	floatx4 gw2[4][4], gw1[4][2];
	for (uint32_t a = 0; a < 4; a++) {
		for (uint32_t b = 0; b < 4; b++) gw2[a][b] = floatx4{ 0.f, 0.f, 0.f, 0.f };
		for (uint32_t b = 0; b < 2; b++) gw1[a][b] = floatx4{ 0.f, 0.f, 0.f, 0.f };
	}
	float gb1[4] = { 0, 0, 0, 0 }, gb2[4] = { 0, 0, 0, 0 }, gw3[4] = { 0, 0, 0, 0 }, gb3 = 0.f, lsum = 0.f;
	const uint32_t npair = (n + 31) / 32;
	for (uint32_t pr = blockIdx.x * 4 + wv; pr < npair; pr += gridDim.x * 4) {
		floatx4 a1[2][4], a2[2][4];
		float dy[2][4];
		for (uint32_t half = 0; half < 2; half++) {
			const uint32_t srow = pr * 32 + half * 16 + r;
			bf16x8 xa;
			if (srow < n) xa = ld8(&X[(size_t) idx[srow] * KP + 8 * g]);
			else for (int q = 0; q < 8; q++) xa[q] = (__bf16) 0.f;
#pragma unroll
			for (uint32_t q = 0; q < 8; q++) B.xt[(8 * g + q) * 32 + half * 16 + r] = xa[q];
			float yr[4];
			forward_tile(s, xa, &B.h1[half * 16 * HD], a1[half], a2[half], yr);
			// dy for rows 4g+q of this half: every lane of the group holds it.
#pragma unroll
			for (uint32_t q = 0; q < 4; q++) {
				const uint32_t row = pr * 32 + half * 16 + 4 * g + q;
				float d = 0.f;
				if (row < n) {
					const float e = yr[q] - tgt[idx[row]];
					d = scale * e;
					if (r == 0) lsum += e * e;
				}
				dy[half][q] = d;
			}
		}
		__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
		__builtin_amdgcn_wave_barrier();
		__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
		// h1 transposed copy and dh2 = dy w3 (z2 > 0).
		for (uint32_t half = 0; half < 2; half++) {
#pragma unroll
			for (uint32_t t = 0; t < 4; t++) {
				const uint32_t col = t * 16 + r;
#pragma unroll
				for (uint32_t q = 0; q < 4; q++) {
					const uint32_t srow = half * 16 + 4 * g + q;
					B.h1t[col * 32 + srow] = B.h1[srow * HD + col];
					const float z2 = a2[half][t][q];
					const float h2 = fmaxf(z2, 0.f);
					gw3[t] += h2 * dy[half][q];
					const float d2 = z2 > 0.f ? dy[half][q] * s.w3[col] : 0.f;
					gb2[t] += d2;
					const __bf16 db = (__bf16) d2;
					B.dh2[srow * HD + col] = db;
					B.dh2t[col * 32 + srow] = db;
				}
			}
			if (r == 0)
				for (uint32_t q = 0; q < 4; q++) gb3 += dy[half][q];
		}
		__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
		__builtin_amdgcn_wave_barrier();
		__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
		// dW2[out][in] += dh2^T h1  (M = out, N = in, K = 32 samples)
#pragma unroll
		for (uint32_t mt = 0; mt < 4; mt++) {
			const bf16x8 A = ld8(&B.dh2t[(mt * 16 + r) * 32 + 8 * g]);
#pragma unroll
			for (uint32_t nt = 0; nt < 4; nt++) gw2[mt][nt] = mfma(A, ld8(&B.h1t[(nt * 16 + r) * 32 + 8 * g]), gw2[mt][nt]);
		}
		// dh1 = (dh2 W2) (z1 > 0)  (M = 16 samples, N = in, K = out 64), per half.
		for (uint32_t half = 0; half < 2; half++) {
			const bf16x8 A0 = ld8(&B.dh2[(half * 16 + r) * HD + 8 * g]);
			const bf16x8 A1 = ld8(&B.dh2[(half * 16 + r) * HD + 32 + 8 * g]);
#pragma unroll
			for (uint32_t nt = 0; nt < 4; nt++) {
				const floatx4 z = { 0.f, 0.f, 0.f, 0.f };
				floatx4 c = mfma(A0, ld8(&s.w2t[(nt * 16 + r) * HD + 8 * g]), z);
				c = mfma(A1, ld8(&s.w2t[(nt * 16 + r) * HD + 32 + 8 * g]), c);
				const uint32_t col = nt * 16 + r;
#pragma unroll
				for (uint32_t q = 0; q < 4; q++) {
					const float d1 = a1[half][nt][q] > 0.f ? c[q] : 0.f;
					gb1[nt] += d1;
					B.dh1t[col * 32 + half * 16 + 4 * g + q] = (__bf16) d1;
				}
			}
		}
		__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
		__builtin_amdgcn_wave_barrier();
		__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
		// dW1[out][in] += dh1^T X  (M = out 64, N = in 32, K = 32 samples)
#pragma unroll
		for (uint32_t mt = 0; mt < 4; mt++) {
			const bf16x8 A = ld8(&B.dh1t[(mt * 16 + r) * 32 + 8 * g]);
#pragma unroll
			for (uint32_t nt = 0; nt < 2; nt++) gw1[mt][nt] = mfma(A, ld8(&B.xt[(nt * 16 + r) * 32 + 8 * g]), gw1[mt][nt]);
		}
		__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
		__builtin_amdgcn_wave_barrier();
		__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
	}
	// Workgroup reduction in LDS, then one partial row per workgroup.
	for (uint32_t mt = 0; mt < 4; mt++)
		for (uint32_t q = 0; q < 4; q++) {
			const uint32_t o = mt * 16 + 4 * g + q;
			for (uint32_t nt = 0; nt < 4; nt++) atomicAdd(&red[OFF_W2 + o * HD + nt * 16 + r], gw2[mt][nt][q]);
			for (uint32_t nt = 0; nt < 2; nt++) atomicAdd(&red[OFF_W1 + o * KP + nt * 16 + r], gw1[mt][nt][q]);
		}
	for (uint32_t t = 0; t < 4; t++) {
		atomicAdd(&red[OFF_B1 + t * 16 + r], gb1[t]);
		atomicAdd(&red[OFF_B2 + t * 16 + r], gb2[t]);
		atomicAdd(&red[OFF_W3 + t * 16 + r], gw3[t]);
	}
	if (r == 0) atomicAdd(&red[OFF_B3], gb3);
	if (lsum != 0.f) atomicAdd(&lred, lsum);
	__syncthreads();
	for (uint32_t q = threadIdx.x; q < NPARAM; q += blockDim.x) partial[(size_t) blockIdx.x * NPARAM + q] = red[q];
	if (threadIdx.x == 0) lpart[blockIdx.x] = lred;
}

// Sums the partials; writes the gradient (and loss) when out != null,
// applies Adam when lr > 0; refreshes the BF16 weight copies.
__global__ void k_mlp_adam(const float *partial, uint32_t nparts, const float *lpart, float *w, float *m, float *v,
                           __bf16 *w1, __bf16 *w2, __bf16 *w2t, float lr, float b1c, float b2c, float *gout,
                           float *loss_out)
{
	const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
	if (p == 0 && loss_out) {
		float L = 0.f;
		for (uint32_t q = 0; q < nparts; q++) L += lpart[q];
		*loss_out = L;
	}
	if (p >= NPARAM) return;
	float gsum = 0.f;
	for (uint32_t q = 0; q < nparts; q++) gsum += partial[(size_t) q * NPARAM + p];
	if (p >= OFF_W1 && p < OFF_B1 && (p % KP) >= NF) gsum = 0.f;   // padding inputs
	if (gout) gout[p] = gsum;
	float wp = w[p];
	if (lr > 0.f) {
		const float mm = 0.9f * m[p] + 0.1f * gsum;
		const float vv = 0.999f * v[p] + 0.001f * gsum * gsum;
		m[p] = mm;
		v[p] = vv;
		wp -= lr * (mm / b1c) / (sqrtf(vv / b2c) + 1e-8f);
		if (p >= OFF_W1 && p < OFF_B1 && (p % KP) >= NF) wp = 0.f;
		w[p] = wp;
	}
	if (p < OFF_B1) w1[p - OFF_W1] = (__bf16) wp;
	if (p >= OFF_W2 && p < OFF_B2) {
		const uint32_t q = p - OFF_W2, o = q / HD, i = q % HD;
		w2[q] = (__bf16) wp;
		w2t[i * HD + o] = (__bf16) wp;
	}
}

// ---------------------------------------------------------------------------
// Features (see oracle/mlp.py for the definitions), one thread per vertex.

struct FeatArgs {
	const uint32_t *fan_off, *fan;
	const uint8_t *closed;
	const double *pos;
	const uint32_t *src;
	uint32_t ns;
	const uint32_t *levels;
	uint32_t lmax, nv;
	double h;
	__bf16 *X;        // [nv][32]
	float *F;         // [nv][10] (optional, FP32 copy)
	float *e;         // [nv] Euclidean distance to the nearest source
};

__device__ void vertex_normal(const FeatArgs &a, uint32_t v, double n[3])
{
	const uint32_t s = a.fan_off[v], c = a.fan_off[v + 1] - s;
	const double *p = a.pos + 3 * (size_t) v;
	n[0] = n[1] = n[2] = 0;
	for (uint32_t k = 0; k + 1 < c; k++) {
		const double *A = a.pos + 3 * (size_t) a.fan[s + k], *Bp = a.pos + 3 * (size_t) a.fan[s + k + 1];
		const double x0 = A[0] - p[0], y0 = A[1] - p[1], z0 = A[2] - p[2];
		const double x1 = Bp[0] - p[0], y1 = Bp[1] - p[1], z1 = Bp[2] - p[2];
		n[0] += y0 * z1 - z0 * y1;
		n[1] += z0 * x1 - x0 * z1;
		n[2] += x0 * y1 - y0 * x1;
	}
	const double nn = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
	if (nn > 0) { n[0] /= nn; n[1] /= nn; n[2] /= nn; }
}

__global__ void k_features(FeatArgs a)
{
	const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
	if (v >= a.nv) return;
	const double *p = a.pos + 3 * (size_t) v;
	uint32_t sbest = a.src[0];
	double e2 = INFINITY;
	for (uint32_t q = 0; q < a.ns; q++) {
		const double *ps = a.pos + 3 * (size_t) a.src[q];
		const double d2 = (p[0] - ps[0]) * (p[0] - ps[0]) + (p[1] - ps[1]) * (p[1] - ps[1]) + (p[2] - ps[2]) * (p[2] - ps[2]);
		if (d2 < e2) { e2 = d2; sbest = a.src[q]; }
	}
	const double e = sqrt(e2);
	const double *ps = a.pos + 3 * (size_t) sbest;
	double nv_[3], ns_[3];
	vertex_normal(a, v, nv_);
	vertex_normal(a, sbest, ns_);
	const uint32_t s = a.fan_off[v], c = a.fan_off[v + 1] - s;
	const uint32_t nl = c ? (a.closed[v] ? c - 1 : c) : 0;
	double lsum = 0, lmin = INFINITY, lmax = 0, cx = 0, cy = 0, cz = 0;
	for (uint32_t k = 0; k < nl; k++) {
		const double *q = a.pos + 3 * (size_t) a.fan[s + k];
		const double d = sqrt((q[0] - p[0]) * (q[0] - p[0]) + (q[1] - p[1]) * (q[1] - p[1]) + (q[2] - p[2]) * (q[2] - p[2]));
		lsum += d; lmin = fmin(lmin, d); lmax = fmax(lmax, d);
		cx += q[0]; cy += q[1]; cz += q[2];
	}
	double f[NF];
	const double lbar = nl ? lsum / nl : 0;
	const double h = a.h;
	double ux = 0, uy = 0, uz = 0;
	if (e > 0) { ux = (p[0] - ps[0]) / e; uy = (p[1] - ps[1]) / e; uz = (p[2] - ps[2]) / e; }
	const uint32_t lv = a.levels[v];
	const double lvd = lv == GP_NIL ? 0.0 : (double) lv;
	f[0] = log1p(e / h);
	f[1] = nv_[0] * ux + nv_[1] * uy + nv_[2] * uz;
	f[2] = (nl && lbar > 0) ? 2 * ((cx / nl - p[0]) * nv_[0] + (cy / nl - p[1]) * nv_[1] + (cz / nl - p[2]) * nv_[2]) / lbar : 0;
	f[3] = nl / 6.0;
	f[4] = lvd * h / (e + h);
	f[5] = ((p[0] - ps[0]) * ns_[0] + (p[1] - ps[1]) * ns_[1] + (p[2] - ps[2]) * ns_[2]) / (e + h);
	f[6] = lbar / h;
	f[7] = nv_[0] * ns_[0] + nv_[1] * ns_[1] + nv_[2] * ns_[2];
	f[8] = lvd / (a.lmax ? a.lmax : 1);
	f[9] = (nl && lmax > 0) ? lmin / lmax : 0;
	__bf16 *x = a.X + (size_t) v * KP;
	for (uint32_t k = 0; k < KP; k++) x[k] = (__bf16) (k < NF ? (float) f[k] : 0.f);
	if (a.F)
		for (uint32_t k = 0; k < NF; k++) a.F[(size_t) v * NF + k] = (float) f[k];
	a.e[v] = (float) e;
}

__global__ void k_to_bf16_rows(const float *F, uint32_t n, __bf16 *X)
{
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	for (uint32_t k = 0; k < KP; k++) X[(size_t) i * KP + k] = (__bf16) (k < NF ? F[(size_t) i * NF + k] : 0.f);
}

__global__ void k_dhat(const float *y, const float *e, uint32_t n, double *dhat)
{
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n) dhat[i] = (double) e[i] * exp((double) y[i]);
}

}  // namespace gp

// ---------------------------------------------------------------------------
// Host side

struct gp_mlp {
	gp::MlpDev d;
	~gp_mlp()
	{
		for (void *p : { (void *) d.w, (void *) d.m, (void *) d.v, (void *) d.w1, (void *) d.w2, (void *) d.w2t })
			(void) hipFree(p);
	}
};

namespace gp {

static int refresh_bf16(gp_mlp *m)
{
	k_mlp_adam<<<(NPARAM + 255) / 256, 256>>>(nullptr, 0, nullptr, m->d.w, m->d.m, m->d.v, m->d.w1, m->d.w2, m->d.w2t,
	                                          0.f, 1.f, 1.f, nullptr, nullptr);
	GP_HIP(hipGetLastError());
	GP_HIP(hipDeviceSynchronize());
	return GP_OK;
}

int mlp_create(uint32_t seed, gp_mlp **out)
{
	*out = nullptr;
	int rc = device_check(nullptr, 0);
	if (rc) return rc;
	auto m = std::make_unique<gp_mlp>();
	if (hipMalloc(&m->d.w, 4 * NPARAM) != hipSuccess || hipMalloc(&m->d.m, 4 * NPARAM) != hipSuccess ||
	    hipMalloc(&m->d.v, 4 * NPARAM) != hipSuccess || hipMalloc(&m->d.w1, 2 * HD * KP) != hipSuccess ||
	    hipMalloc(&m->d.w2, 2 * HD * HD) != hipSuccess || hipMalloc(&m->d.w2t, 2 * HD * HD) != hipSuccess)
		return fail(GP_ERR_NOMEM, "mlp: allocation failed");
	// He initialisation (normal), zero biases, zero padding columns.
	std::vector<float> w(NPARAM, 0.f);
	std::mt19937_64 rng(seed);
	std::normal_distribution<double> nd(0.0, 1.0);
	for (uint32_t o = 0; o < HD; o++)
		for (uint32_t i = 0; i < NF; i++) w[OFF_W1 + o * KP + i] = (float) (nd(rng) * std::sqrt(2.0 / NF));
	for (uint32_t q = 0; q < HD * HD; q++) w[OFF_W2 + q] = (float) (nd(rng) * std::sqrt(2.0 / HD));
	for (uint32_t q = 0; q < HD; q++) w[OFF_W3 + q] = (float) (nd(rng) * std::sqrt(1.0 / HD));
	GP_HIP(hipMemcpy(m->d.w, w.data(), 4 * NPARAM, hipMemcpyHostToDevice));
	std::vector<float> zero(NPARAM, 0.f);
	GP_HIP(hipMemcpy(m->d.m, zero.data(), 4 * NPARAM, hipMemcpyHostToDevice));
	GP_HIP(hipMemcpy(m->d.v, zero.data(), 4 * NPARAM, hipMemcpyHostToDevice));
	rc = refresh_bf16(m.get());
	if (rc) return rc;
	*out = m.release();
	return GP_OK;
}

int mlp_params(gp_mlp *m, float *get, const float *set, uint32_t n)
{
	if (n != NPARAM) return fail(GP_ERR_CAPACITY, "mlp: the parameter vector has %u entries", NPARAM);
	if (get) GP_HIP(hipMemcpy(get, m->d.w, 4 * NPARAM, hipMemcpyDeviceToHost));
	if (set) {
		GP_HIP(hipMemcpy(m->d.w, set, 4 * NPARAM, hipMemcpyHostToDevice));
		return refresh_bf16(m);
	}
	return GP_OK;
}

struct DevBuf {
	void *p = nullptr;
	~DevBuf() { if (p) (void) hipFree(p); }
	int alloc(size_t b)
	{
		if (hipMalloc(&p, std::max<size_t>(b, 4)) != hipSuccess) return fail(GP_ERR_NOMEM, "device allocation failed");
		return GP_OK;
	}
};

// Sources and BFS levels in device memory, with the largest level (the
// inputs of the feature kernel; the levels are the toplesets of the sources).
struct FeatInputs {
	DevBuf src, lev;
	uint32_t ns = 0, lmax = 0;
	int load(Che &mesh, const uint32_t *s, uint32_t n, const uint32_t *levels)
	{
		int rc = upload_mesh(mesh);
		if (rc) return rc;
		if ((rc = src.alloc(4ull * n)) || (rc = lev.alloc(4ull * mesh.nv))) return rc;
		GP_HIP(hipMemcpy(src.p, s, 4ull * n, hipMemcpyHostToDevice));
		GP_HIP(hipMemcpy(lev.p, levels, 4ull * mesh.nv, hipMemcpyHostToDevice));
		ns = n;
		lmax = 0;
		for (uint32_t v = 0; v < mesh.nv; v++)
			if (levels[v] != NIL) lmax = std::max(lmax, levels[v]);
		return GP_OK;
	}
};

static void launch_features(Che &mesh, const FeatInputs &in, __bf16 *X, float *F, float *e)
{
	FeatArgs a{ mesh.dev->fan_off, mesh.dev->fan, mesh.dev->closed, mesh.dev->pos, (const uint32_t *) in.src.p, in.ns,
	            (const uint32_t *) in.lev.p, in.lmax, mesh.nv, mesh.mean_edge, X, F, e };
	k_features<<<(mesh.nv + 255) / 256, 256>>>(a);
}

// Features of every vertex on the GPU (levels: host toplesets of the sources).
int mlp_features_dev(Che &mesh, const uint32_t *src, uint32_t ns, const uint32_t *levels, __bf16 *X, float *F,
                     float *e)
{
	FeatInputs in;
	int rc = in.load(mesh, src, ns, levels);
	if (rc) return rc;
	launch_features(mesh, in, X, F, e);
	GP_HIP(hipGetLastError());
	GP_HIP(hipDeviceSynchronize());
	return GP_OK;
}

int mlp_features(Che &mesh, const uint32_t *src, uint32_t ns, const uint32_t *levels, float *F_out, float *e_out)
{
	DevBuf X, F, e;
	int rc;
	if ((rc = X.alloc(2ull * KP * mesh.nv)) || (rc = F.alloc(4ull * NF * mesh.nv)) || (rc = e.alloc(4ull * mesh.nv))) return rc;
	rc = mlp_features_dev(mesh, src, ns, levels, (__bf16 *) X.p, (float *) F.p, (float *) e.p);
	if (rc) return rc;
	if (F_out) GP_HIP(hipMemcpy(F_out, F.p, 4ull * NF * mesh.nv, hipMemcpyDeviceToHost));
	if (e_out) GP_HIP(hipMemcpy(e_out, e.p, 4ull * mesh.nv, hipMemcpyDeviceToHost));
	return GP_OK;
}

int mlp_predict(gp_mlp *m, Che &mesh, const uint32_t *src, uint32_t ns, const uint32_t *levels, double *dhat,
                float *ms_out)
{
	DevBuf X, e, y, d;
	int rc;
	const uint32_t nv = mesh.nv;
	if ((rc = X.alloc(2ull * KP * nv)) || (rc = e.alloc(4ull * nv)) || (rc = y.alloc(4ull * nv)) || (rc = d.alloc(8ull * nv)))
		return rc;
	FeatInputs in;
	if ((rc = in.load(mesh, src, ns, levels))) return rc;
	// Device time of the prediction proper: features and network on
	// device-resident inputs (the levels are the sources' toplesets, which
	// every schedule computes before its timed sweeps).
	hipEvent_t e0, e1;
	GP_HIP(hipEventCreate(&e0));
	GP_HIP(hipEventCreate(&e1));
	GP_HIP(hipEventRecord(e0));
	launch_features(mesh, in, (__bf16 *) X.p, nullptr, (float *) e.p);
	const uint32_t ntile = (nv + 15) / 16;
	const uint32_t grid = std::max<uint32_t>(1, std::min<uint32_t>(2048, (ntile + 3) / 4));
	k_mlp_forward<<<grid, 256>>>((const __bf16 *) X.p, nv, m->d.w, m->d.w1, m->d.w2, (float *) y.p);
	k_dhat<<<(nv + 255) / 256, 256>>>((const float *) y.p, (const float *) e.p, nv, (double *) d.p);
	GP_HIP(hipGetLastError());
	GP_HIP(hipEventRecord(e1));
	GP_HIP(hipEventSynchronize(e1));
	float ms = 0;
	GP_HIP(hipEventElapsedTime(&ms, e0, e1));
	if (ms_out) *ms_out = ms;
	(void) hipEventDestroy(e0);
	(void) hipEventDestroy(e1);
	GP_HIP(hipMemcpy(dhat, d.p, 8ull * nv, hipMemcpyDeviceToHost));
	return GP_OK;
}

// The same prediction on device-resident inputs: the sources and their BFS
// levels in device memory (lmax: the largest level), dhat written to device
// memory, every kernel on `stream` -- the same kernels and inputs as
// mlp_predict, hence the same dhat, without the host round trips.
int mlp_predict_device(gp_mlp *m, Che &mesh, const uint32_t *d_src, uint32_t ns, const uint32_t *d_levels,
                       uint32_t lmax, double *d_dhat, void *stream_)
{
	const hipStream_t stream = (hipStream_t) stream_;
	int rc = upload_mesh(mesh);
	if (rc) return rc;
	DevBuf X, e, y;
	const uint32_t nv = mesh.nv;
	if ((rc = X.alloc(2ull * KP * nv)) || (rc = e.alloc(4ull * nv)) || (rc = y.alloc(4ull * nv))) return rc;
	FeatArgs a{ mesh.dev->fan_off, mesh.dev->fan, mesh.dev->closed, mesh.dev->pos, d_src, ns, d_levels, lmax, nv,
	            mesh.mean_edge, (__bf16 *) X.p, nullptr, (float *) e.p };
	k_features<<<(nv + 255) / 256, 256, 0, stream>>>(a);
	const uint32_t ntile = (nv + 15) / 16;
	const uint32_t grid = std::max<uint32_t>(1, std::min<uint32_t>(2048, (ntile + 3) / 4));
	k_mlp_forward<<<grid, 256, 0, stream>>>((const __bf16 *) X.p, nv, m->d.w, m->d.w1, m->d.w2, (float *) y.p);
	k_dhat<<<(nv + 255) / 256, 256, 0, stream>>>((const float *) y.p, (const float *) e.p, nv, d_dhat);
	GP_HIP(hipGetLastError());
	GP_HIP(hipStreamSynchronize(stream));   // before the scratch is freed
	return GP_OK;
}

// Forward on caller features F[n][10]: y[n].
int mlp_forward(gp_mlp *m, const float *F, uint32_t n, float *y)
{
	DevBuf dF, X, dy;
	int rc;
	if ((rc = dF.alloc(4ull * NF * n)) || (rc = X.alloc(2ull * KP * n)) || (rc = dy.alloc(4ull * n))) return rc;
	GP_HIP(hipMemcpy(dF.p, F, 4ull * NF * n, hipMemcpyHostToDevice));
	k_to_bf16_rows<<<(n + 255) / 256, 256>>>((const float *) dF.p, n, (__bf16 *) X.p);
	const uint32_t ntile = (n + 15) / 16;
	k_mlp_forward<<<std::max<uint32_t>(1, std::min<uint32_t>(2048, (ntile + 3) / 4)), 256>>>(
		(const __bf16 *) X.p, n, m->d.w, m->d.w1, m->d.w2, (float *) dy.p);
	GP_HIP(hipGetLastError());
	GP_HIP(hipMemcpy(y, dy.p, 4ull * n, hipMemcpyDeviceToHost));
	return GP_OK;
}

// Training on caller data.  steps == 0: one full-batch gradient into grad_out
// (no update).  Otherwise `steps` Adam steps on minibatches of `batch`
// samples drawn (without replacement within an epoch) with the given seed.
int mlp_train(gp_mlp *m, const float *F, const float *t, uint32_t n, uint32_t steps, uint32_t batch, float lr,
              uint32_t seed, float *grad_out, float *loss_out)
{
	if (n == 0) return fail(GP_ERR_INVALID, "mlp_train: no samples");
	DevBuf dF, X, dt, didx, part, lpart, dg, dl;
	int rc;
	const uint32_t grid = 256;
	if ((rc = dF.alloc(4ull * NF * n)) || (rc = X.alloc(2ull * KP * n)) || (rc = dt.alloc(4ull * n)) ||
	    (rc = didx.alloc(4ull * n)) || (rc = part.alloc(4ull * NPARAM * grid)) || (rc = lpart.alloc(4ull * grid)) ||
	    (rc = dg.alloc(4ull * NPARAM)) || (rc = dl.alloc(4)))
		return rc;
	GP_HIP(hipMemcpy(dF.p, F, 4ull * NF * n, hipMemcpyHostToDevice));
	GP_HIP(hipMemcpy(dt.p, t, 4ull * n, hipMemcpyHostToDevice));
	k_to_bf16_rows<<<(n + 255) / 256, 256>>>((const float *) dF.p, n, (__bf16 *) X.p);
	GP_HIP(hipGetLastError());
	std::vector<uint32_t> perm(n);
	for (uint32_t q = 0; q < n; q++) perm[q] = q;
	if (steps == 0) {
		GP_HIP(hipMemcpy(didx.p, perm.data(), 4ull * n, hipMemcpyHostToDevice));
		k_mlp_grad<<<grid, 256>>>((const __bf16 *) X.p, (const float *) dt.p, (const uint32_t *) didx.p, n, 2.0f / n,
		                          m->d.w, m->d.w1, m->d.w2, m->d.w2t, (float *) part.p, (float *) lpart.p);
		k_mlp_adam<<<(NPARAM + 255) / 256, 256>>>((const float *) part.p, grid, (const float *) lpart.p, m->d.w, m->d.m,
		                                          m->d.v, m->d.w1, m->d.w2, m->d.w2t, 0.f, 1.f, 1.f, (float *) dg.p,
		                                          (float *) dl.p);
		GP_HIP(hipGetLastError());
		if (grad_out) GP_HIP(hipMemcpy(grad_out, dg.p, 4ull * NPARAM, hipMemcpyDeviceToHost));
		float L = 0;
		GP_HIP(hipMemcpy(&L, dl.p, 4, hipMemcpyDeviceToHost));
		if (loss_out) *loss_out = L / n;
		return GP_OK;
	}
	std::mt19937_64 rng(seed);
	const uint32_t bs = std::min(batch ? batch : 8192u, n);
	uint32_t cursor = n;
	float last = 0;
	for (uint32_t s = 0; s < steps; s++) {
		if (cursor + bs > n) {
			std::shuffle(perm.begin(), perm.end(), rng);
			GP_HIP(hipMemcpy(didx.p, perm.data(), 4ull * n, hipMemcpyHostToDevice));
			cursor = 0;
		}
		m->d.step++;
		const float b1c = 1.f - std::pow(0.9f, (float) m->d.step), b2c = 1.f - std::pow(0.999f, (float) m->d.step);
		const float lr_s = lr * 0.5f * (1.f + std::cos(3.14159265f * s / steps));   // cosine decay
		const uint32_t g = std::max<uint32_t>(1, std::min<uint32_t>(grid, (bs + 127) / 128));
		k_mlp_grad<<<g, 256>>>((const __bf16 *) X.p, (const float *) dt.p, (const uint32_t *) didx.p + cursor, bs,
		                       2.0f / bs, m->d.w, m->d.w1, m->d.w2, m->d.w2t, (float *) part.p, (float *) lpart.p);
		k_mlp_adam<<<(NPARAM + 255) / 256, 256>>>((const float *) part.p, g, (const float *) lpart.p, m->d.w, m->d.m,
		                                          m->d.v, m->d.w1, m->d.w2, m->d.w2t, lr_s, b1c, b2c, nullptr,
		                                          (float *) dl.p);
		GP_HIP(hipGetLastError());
		cursor += bs;
		if (s + 1 == steps) {
			GP_HIP(hipMemcpy(&last, dl.p, 4, hipMemcpyDeviceToHost));
			last /= bs;
		}
	}
	GP_HIP(hipDeviceSynchronize());
	if (loss_out) *loss_out = last;
	return GP_OK;
}

}  // namespace gp

extern "C" void gp_mlp_destroy(gp_mlp *m) { delete m; }
