// Device rings: frontier BFS levels and a counting sort by level (see levels.hpp).
#include "levels.hpp"

#include <algorithm>
#include <memory>
#include <type_traits>

#include "gp_internal.hpp"

namespace gp {

// Frontier BFS state: launch k expands the frontier queue[fa, fb) of level L
// (slot (k-1) & 3 advanced by the count cnt[(k-1) & 3] of the vertices the
// previous launch appended), appending the vertices it lowers to level L + 1
// after fb through cnt[k & 3].
struct LSlot {
	uint32_t L, fa, fb, done, fresh, pad0, pad1, pad2;
};
struct LevelState {
	LSlot slot[4];
	uint32_t cnt[4];
	uint32_t done, launches, max_level, pad;
};

__global__ void k_lv_zero(uint32_t *p, uint32_t n)
{
	for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) p[v] = 0;
}

// Levels before the first round: unreached everywhere.
__global__ void k_lv_clear(uint32_t *level, uint32_t n)
{
	for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) level[v] = GP_NIL;
}

// Seeds the frontier BFS: sources src[first, ns) get level 0 and form the
// frontier.
__global__ void k_lv_seed(const uint32_t *src, uint32_t first, uint32_t ns, uint32_t *level, uint32_t *queue,
                          LevelState *ls)
{
	for (uint32_t q = first + threadIdx.x; q < ns; q += blockDim.x) {
		level[src[q]] = 0;
		queue[q - first] = src[q];
	}
	if (threadIdx.x == 0) {
		uint32_t *w = (uint32_t *) ls;
		for (uint32_t q = 0; q < sizeof(LevelState) / 4; q++) w[q] = 0;
		ls->slot[3].L = 0; ls->slot[3].fa = 0; ls->slot[3].fb = ns - first; ls->slot[3].fresh = 1;
		ls->slot[3].done = ns > first ? 0u : 1u;
		ls->done = ls->slot[3].done;
	}
}

// One BFS level over all workgroups.  A neighbour whose level drops to L + 1
// (atomicMin: exactly one thread sees it drop) is appended once.  A lane
// issues the atomics of LVU fan entries back to back (slots past its fan
// apply the no-op minimum NIL to its own vertex), so one level costs one
// atomic round trip per LVU entries, not one per entry; the appends of a wave
// take one counter atomic per pass (ballot per entry slot).
constexpr uint32_t LVU = 8;

__global__ __launch_bounds__(256) void k_lv_step(const uint32_t *fan_off, const uint32_t *fan, uint32_t *level,
                                                 uint32_t *queue, LevelState *ls, uint32_t k4)
{
	const uint32_t kp = (k4 + 3u) & 3u;
	const LSlot prev = ls->slot[kp];
	const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
	if (prev.done) {
		if (lead) ls->slot[k4] = prev;
		return;
	}
	LSlot cur = prev;
	cur.fresh = 0;
	if (!prev.fresh) {
		const uint32_t c = ls->cnt[kp];
		cur.L = prev.L + 1; cur.fa = prev.fb; cur.fb = prev.fb + c; cur.done = c == 0;
	}
	if (lead) {
		ls->slot[k4] = cur;
		ls->cnt[(k4 + 1u) & 3u] = 0;
		ls->launches = ls->launches + 1;
		if (cur.done) ls->done = 1;
	}
	if (cur.done) return;
	const uint32_t lane = threadIdx.x & 63u;
	const unsigned long long below = (1ull << lane) - 1ull;
	const uint32_t nl1 = cur.L + 1;
	const uint32_t stride = gridDim.x * blockDim.x;
	// Wave-uniform loops: every lane of a wave runs the same iterations (the
	// ballots and the shuffle below need them all).
	for (uint32_t b = cur.fa + (blockIdx.x * blockDim.x + (threadIdx.x & ~63u)); b < cur.fb; b += stride) {
		const uint32_t idx = b + lane;
		const bool act = idx < cur.fb;
		const uint32_t u = act ? queue[idx] : 0u;
		const uint32_t e0 = act ? fan_off[u] : 0u, e1 = act ? fan_off[u + 1] : 0u;
		const uint32_t len = e1 - e0;
		uint32_t mx = len;
		for (uint32_t o = 32; o > 0; o >>= 1) mx = max(mx, (uint32_t) __shfl_xor(mx, o));
		for (uint32_t eb = 0; eb < mx; eb += LVU) {
			uint32_t w[LVU];
			uint32_t app = 0;
			if (eb < len) {
#pragma unroll
				for (uint32_t k = 0; k < LVU; k++) w[k] = fan[min(e0 + eb + k, e1 - 1u)];
#pragma unroll
				for (uint32_t k = 0; k < LVU; k++) {
					const bool ok = eb + k < len;
					const uint32_t old = atomicMin(&level[ok ? w[k] : u], ok ? nl1 : GP_NIL);
					app |= (ok && old > nl1) ? (1u << k) : 0u;
				}
			}
			unsigned long long m[LVU];
			uint32_t tot = 0;
#pragma unroll
			for (uint32_t k = 0; k < LVU; k++) {
				m[k] = __ballot((app >> k) & 1u);
				tot += (uint32_t) __popcll(m[k]);
			}
			if (tot) {
				uint32_t base = 0;
				if (lane == 0) base = atomicAdd(&ls->cnt[k4], tot);
				base = cur.fb + __shfl(base, 0);
#pragma unroll
				for (uint32_t k = 0; k < LVU; k++) {
					if ((app >> k) & 1u) queue[base + (uint32_t) __popcll(m[k] & below)] = w[k];
					base += (uint32_t) __popcll(m[k]);
				}
			}
		}
	}
}

// Counting sort of the reached vertices by level and, inside a level, by slot
// count (longest fans first), so that a 64-position chunk of a ring holds fans
// of one length: histogram of the keys (and the largest level), exclusive scan
// into the cursors and limits (one workgroup), scatter into sorted[].  The
// order inside a level is free (a sweep relaxes whole rings from the previous
// buffer).  Global atomics on a few thousand counters serialise, so each
// workgroup counts a block of SORT_BLOCK vertices in LDS bins (keys lo .. lo +
// SORT_BINS - 1 of the block; keys past them, rare, count globally) and adds
// or reserves one global count per key present in the block.
constexpr uint32_t SORT_BLOCK = 8192, SORT_BINS = 12288, SORT_T = 1024;

// Sort key of vertex v: level * KS + slot class (KS or more slots first, then
// KS - 1, ..., 1), NIL when unreached.
__device__ __forceinline__ uint32_t sort_key(const uint32_t *level, const uint32_t *fan_off, const uint8_t *closed,
                                             uint32_t v)
{
	const uint32_t l = level[v];
	if (l == GP_NIL) return GP_NIL;
	const uint32_t n = fan_off[v + 1] - fan_off[v];
	const uint32_t sl = (n && closed[v]) ? n - 1 : n;
	return l * LEVEL_KEYS + (LEVEL_KEYS - min(max(sl, 1u), LEVEL_KEYS));
}

struct KeySrc {
	const uint32_t *level, *fan_off;
	const uint8_t *closed;
	__device__ uint32_t operator()(uint32_t v) const { return sort_key(level, fan_off, closed, v); }
};

// Smallest and largest key of the block [b0, b1) (every thread).
__device__ void block_key_range(const KeySrc &key, uint32_t b0, uint32_t b1, uint32_t *red, uint32_t &lo,
                                uint32_t &hi)
{
	lo = GP_NIL; hi = 0;
	for (uint32_t v = b0 + threadIdx.x; v < b1; v += SORT_T) {
		const uint32_t k = key(v);
		if (k != GP_NIL) { lo = min(lo, k); hi = max(hi, k); }
	}
	for (uint32_t o = 32; o > 0; o >>= 1) {
		lo = min(lo, (uint32_t) __shfl_xor(lo, o));
		hi = max(hi, (uint32_t) __shfl_xor(hi, o));
	}
	const uint32_t w = threadIdx.x >> 6;
	if ((threadIdx.x & 63u) == 0) { red[w] = lo; red[16 + w] = hi; }
	__syncthreads();
	lo = GP_NIL; hi = 0;
	for (uint32_t q = 0; q < SORT_T / 64; q++) { lo = min(lo, red[q]); hi = max(hi, red[16 + q]); }
}

// LDS counts of the block's keys lo .. lo + SORT_BINS - 1; the others are
// counted in glob (when given).
__device__ void block_key_counts(const KeySrc &key, uint32_t b0, uint32_t b1, uint32_t lo, uint32_t *bins,
                                 uint32_t *glob)
{
	for (uint32_t q = threadIdx.x; q < SORT_BINS; q += SORT_T) bins[q] = 0;
	__syncthreads();
	for (uint32_t v = b0 + threadIdx.x; v < b1; v += SORT_T) {
		const uint32_t k = key(v);
		if (k == GP_NIL) continue;
		if (k - lo < SORT_BINS) atomicAdd(&bins[k - lo], 1u);
		else if (glob) atomicAdd(&glob[k], 1u);
	}
	__syncthreads();
}

__global__ __launch_bounds__(SORT_T) void k_lv_hist(KeySrc key, uint32_t n, uint32_t *hist, LevelState *ls)
{
	__shared__ uint32_t bins[SORT_BINS];
	__shared__ uint32_t red[32];
	const uint32_t b0 = blockIdx.x * SORT_BLOCK, b1 = min(n, b0 + SORT_BLOCK);
	uint32_t lo, hi;
	block_key_range(key, b0, b1, red, lo, hi);
	if (lo == GP_NIL) return;   // nothing reached in the block (uniform)
	block_key_counts(key, b0, b1, lo, bins, hist);
	const uint32_t range = min(SORT_BINS, hi - lo + 1);
	for (uint32_t q = threadIdx.x; q < range; q += SORT_T)
		if (bins[q]) atomicAdd(&hist[lo + q], bins[q]);
	if (threadIdx.x == 0) atomicMax(&ls->max_level, hi / LEVEL_KEYS);
}

// Exclusive scan of the key counts into the scatter cursors; limits[l] = the
// first position of level l, limits[max_level + 1] = the reached count.
__global__ __launch_bounds__(1024) void k_lv_scan(uint32_t *hist, uint32_t *limits, uint32_t *nl, const LevelState *ls)
{
	__shared__ uint32_t part[1024];
	const uint32_t nlev = ls->max_level + 1, nb = nlev * LEVEL_KEYS;
	uint32_t carry = 0;
	for (uint32_t b0 = 0; b0 < nb; b0 += 1024) {
		const uint32_t q = b0 + threadIdx.x;
		const uint32_t x = q < nb ? hist[q] : 0u;
		part[threadIdx.x] = x;
		__syncthreads();
		for (uint32_t o = 1; o < 1024; o <<= 1) {
			const uint32_t y = threadIdx.x >= o ? part[threadIdx.x - o] : 0u;
			__syncthreads();
			part[threadIdx.x] += y;
			__syncthreads();
		}
		const uint32_t excl = part[threadIdx.x] + carry - x;
		if (q < nb) {
			hist[q] = excl;   // becomes the scatter cursor of key q
			if (q % LEVEL_KEYS == 0) limits[q / LEVEL_KEYS] = excl;
		}
		const uint32_t tot = part[1023];
		__syncthreads();
		carry += tot;
	}
	if (threadIdx.x == 0) { limits[nlev] = carry; *nl = nlev + 1; }
}

__global__ __launch_bounds__(SORT_T) void k_lv_scatter(KeySrc key, uint32_t n, uint32_t *cursor, uint32_t *sorted)
{
	__shared__ uint32_t bins[SORT_BINS];
	__shared__ uint32_t red[32];
	const uint32_t b0 = blockIdx.x * SORT_BLOCK, b1 = min(n, b0 + SORT_BLOCK);
	uint32_t lo, hi;
	block_key_range(key, b0, b1, red, lo, hi);
	if (lo == GP_NIL) return;
	block_key_counts(key, b0, b1, lo, bins, nullptr);
	// Reserve the block's run of each key; the bins become its cursors.
	const uint32_t range = min(SORT_BINS, hi - lo + 1);
	for (uint32_t q = threadIdx.x; q < range; q += SORT_T)
		if (bins[q]) bins[q] = atomicAdd(&cursor[lo + q], bins[q]);
	__syncthreads();
	for (uint32_t v = b0 + threadIdx.x; v < b1; v += SORT_T) {
		const uint32_t k = key(v);
		if (k == GP_NIL) continue;
		sorted[k - lo < SORT_BINS ? atomicAdd(&bins[k - lo], 1u) : atomicAdd(&cursor[k], 1u)] = v;
	}
}

// Slot rows of each 64-position chunk of sorted[] (its longest fan; 0 past the
// reached vertices), one wave per chunk.
__global__ void k_lv_chunk_slots(const uint32_t *sorted, const uint32_t *limits, const uint32_t *nl,
                                 const uint32_t *fan_off, const uint8_t *closed, uint32_t nchunks, uint32_t *cslots)
{
	const uint32_t nr = limits[*nl - 1];
	const uint32_t lane = threadIdx.x & 63u;
	const uint32_t nw = (gridDim.x * blockDim.x) >> 6;
	for (uint32_t c = (blockIdx.x * blockDim.x + threadIdx.x) >> 6; c < nchunks; c += nw) {
		const uint32_t p = (c << 6) + lane;
		uint32_t sl = 0;
		if (p < nr) {
			const uint32_t v = sorted[p];
			const uint32_t n = fan_off[v + 1] - fan_off[v];
			sl = (n && closed[v]) ? n - 1 : n;
		}
		for (uint32_t o = 32; o > 0; o >>= 1) sl = max(sl, (uint32_t) __shfl_xor(sl, o));
		if (lane == 0) cslots[c] = sl;
	}
}

// out[q] = in[0] + ... + in[q - 1] for q <= n (one workgroup).
__global__ __launch_bounds__(1024) void k_exclusive_scan(const uint32_t *in, uint32_t n, uint32_t *out)
{
	__shared__ uint32_t part[1024];
	uint32_t carry = 0;
	for (uint32_t b0 = 0; b0 < n; b0 += 1024) {
		const uint32_t q = b0 + threadIdx.x;
		const uint32_t x = q < n ? in[q] : 0u;
		part[threadIdx.x] = x;
		__syncthreads();
		for (uint32_t o = 1; o < 1024; o <<= 1) {
			const uint32_t y = threadIdx.x >= o ? part[threadIdx.x - o] : 0u;
			__syncthreads();
			part[threadIdx.x] += y;
			__syncthreads();
		}
		if (q < n) out[q] = part[threadIdx.x] + carry - x;
		const uint32_t tot = part[1023];
		__syncthreads();
		carry += tot;
	}
	if (threadIdx.x == 0) out[n] = carry;
}

constexpr uint32_t LV = 32;   // level launches per graph replay

Levels::~Levels()
{
	for (void *p : { (void *) level, (void *) queue, (void *) hist, (void *) sorted, (void *) limits, (void *) nl,
	                 (void *) ls, (void *) cslots, (void *) coff })
		(void) hipFree(p);
	if (h_ls) (void) hipHostFree(h_ls);
	if (graph) (void) hipGraphExecDestroy(graph);
	for (hipEvent_t e : ev)
		if (e) (void) hipEventDestroy(e);
}

int Levels::create(const DeviceMesh *mesh, hipStream_t s)
{
	dm = mesh;
	nv = mesh->nv;
	stream = s;
	if (hipMalloc(&level, 4ull * nv) != hipSuccess || hipMalloc(&queue, 4ull * nv) != hipSuccess ||
	    hipMalloc(&hist, 4ull * ((size_t) nv * LEVEL_KEYS + 1)) != hipSuccess ||
	    hipMalloc(&sorted, 4ull * nv) != hipSuccess || hipMalloc(&cslots, 4ull * (nchunks() + 1)) != hipSuccess ||
	    hipMalloc(&coff, 4ull * (nchunks() + 1)) != hipSuccess ||
	    hipMalloc(&limits, 4ull * (nv + 2)) != hipSuccess || hipMalloc(&nl, 4) != hipSuccess ||
	    hipMalloc(&ls, sizeof(LevelState)) != hipSuccess || hipHostMalloc(&h_ls, 2 * sizeof(LevelState)) != hipSuccess)
		return fail(GP_ERR_NOMEM, "levels: device allocation failed");
	GP_HIP(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
	GP_HIP(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
	// Enough workgroups for a frontier of 64k vertices per pass.
	// One wave per workgroup, as the sweeps.
	const uint32_t grid = std::max<uint32_t>(1, std::min<uint32_t>(1024, (nv + 63) / 64));
	lv_grid = grid;
	if (!graphs_enabled()) return GP_OK;
	const uint32_t *fo = dm->fan_off, *fa = dm->fan;
	const uint32_t k4s[4] = { 0, 1, 2, 3 };
	GraphChain gc;
	for (uint32_t q = 0; q < LV; q++) {
		void *args[] = { (void *) &fo, (void *) &fa, (void *) &level, (void *) &queue, (void *) &ls,
		                 (void *) &k4s[q & 3u] };
		gc.add((const void *) k_lv_step, dim3(grid), dim3(64), args);
	}
	const hipError_t e = gc.instantiate(&graph);
	if (e != hipSuccess) return fail(GP_ERR_HIP, "levels graph: %s", hipGetErrorString(e));
	return GP_OK;
}

static uint32_t fill_grid(uint32_t n) { return std::max<uint32_t>(1, std::min<uint32_t>(2048, (n + 255) / 256)); }

void Levels::clear() { k_lv_clear<<<fill_grid(nv), 256, 0, stream>>>(level, nv); }

void Levels::seed(const uint32_t *d_src, uint32_t first, uint32_t ns)
{
	k_lv_seed<<<1, 256, 0, stream>>>(d_src, first, ns, level, queue, ls);
}

// Replays until the state reports done, two in flight: each replay is
// followed by a kernel that writes the state into its pinned snapshot and an
// event; the host waits for the older replay while the newer one runs (a
// replay queued behind the one that finishes meets done slots and returns at
// once).
int Levels::expand()
{
	GP_HIP(hipGetLastError());
	auto replay = [&](uint32_t s) -> int {
		if (graph) {
			GP_HIP(hipGraphLaunch(graph, stream));
		} else {
			for (uint32_t q = 0; q < LV; q++)
				k_lv_step<<<lv_grid, 64, 0, stream>>>(dm->fan_off, dm->fan, level, queue, ls, q & 3u);
		}
		snapshot(ls, h_ls + s, stream);
		GP_HIP(hipGetLastError());
		GP_HIP(hipEventRecord(ev[s], stream));
		return GP_OK;
	};
	int rc = replay(0);
	if (rc) return rc;
	for (uint32_t reps = 0, cur = 0;; reps++, cur ^= 1u) {
		// Every replay but an empty one lowers at least one vertex.
		if (reps > nv / LV + 2) return fail(GP_ERR_KERNEL, "levels: BFS did not finish");
		if ((rc = replay(cur ^ 1u))) return rc;
		GP_HIP(hipEventSynchronize(ev[cur]));
		if (h_ls[cur].done) break;
	}
	GP_HIP(hipStreamSynchronize(stream));
	return GP_OK;
}

void Levels::sort()
{
	const uint32_t nk = nv * LEVEL_KEYS + 1;
	const uint32_t g = fill_grid(nk), gb = (nv + SORT_BLOCK - 1) / SORT_BLOCK;
	const KeySrc key{ level, dm->fan_off, dm->closed };
	k_lv_zero<<<g, 256, 0, stream>>>(hist, nk);
	k_lv_hist<<<gb, SORT_T, 0, stream>>>(key, nv, hist, ls);
	k_lv_scan<<<1, 1024, 0, stream>>>(hist, limits, nl, ls);
	k_lv_scatter<<<gb, SORT_T, 0, stream>>>(key, nv, hist, sorted);
	const uint32_t nc = nchunks();
	k_lv_chunk_slots<<<fill_grid(64 * nc), 256, 0, stream>>>(sorted, limits, nl, dm->fan_off, dm->closed, nc, cslots);
	k_exclusive_scan<<<1, 1024, 0, stream>>>(cslots, nc, coff);
}

// BFS levels of the sources (NIL where unreached) on the device, copied into
// levels[nv]: the same sets as the host toplesets' levels.
int levels_gpu(Che &m, const uint32_t *src, uint32_t ns, uint32_t *levels)
{
	if (!ns) return fail(GP_ERR_INVALID, "levels: no sources");
	for (uint32_t q = 0; q < ns; q++)
		if (src[q] >= m.nv) return fail(GP_ERR_INVALID, "levels: source %u out of range", src[q]);
	int rc = upload_mesh(m);
	if (rc) return rc;
	hipStream_t stream = nullptr;
	GP_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
	std::unique_ptr<std::remove_pointer<hipStream_t>::type, hipError_t (*)(hipStream_t)> sg(stream, hipStreamDestroy);
	uint32_t *d_src = nullptr;
	if (hipMalloc(&d_src, 4ull * ns) != hipSuccess) return fail(GP_ERR_NOMEM, "levels: device allocation failed");
	std::unique_ptr<uint32_t, hipError_t (*)(void *)> dg(d_src, hipFree);
	GP_HIP(hipMemcpy(d_src, src, 4ull * ns, hipMemcpyHostToDevice));
	Levels lv;
	if ((rc = lv.create(m.dev, stream))) return rc;
	lv.clear();
	lv.seed(d_src, 0, ns);
	if ((rc = lv.expand())) return rc;
	GP_HIP(hipMemcpyAsync(levels, lv.level, 4ull * m.nv, hipMemcpyDeviceToHost, stream));
	GP_HIP(hipStreamSynchronize(stream));
	return GP_OK;
}

}  // namespace gp
