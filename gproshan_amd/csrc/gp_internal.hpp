// Internal declarations of libgproshan_ptp (host side).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/gproshan_ptp.h"

namespace gp {

constexpr uint32_t NIL = GP_NIL;

// Sets the thread-local error text and returns the status.
int fail(int status, const char *fmt, ...);

inline uint32_t he_next(uint32_t he) { return 3u * (he / 3u) + (he + 1u) % 3u; }
inline uint32_t he_prev(uint32_t he) { return 3u * (he / 3u) + (he + 2u) % 3u; }

struct DeviceMesh;  // ptp_gpu.hip

// Compact half-edge mesh plus the per-vertex triangle fans the solvers use.
//
// Fan of v: the star walk from EVT[v] (he <- OT[prev(he)], include/che.h:10)
// visits half-edges he_0..he_{m-1} leaving v; the fan list is
// L = (VT[next(he_0)], ..., VT[next(he_{m-1})], VT[prev(he_{m-1})]), m+1 ids,
// and triangle k of the star is (v, L[k], L[k+1]).  For a closed fan
// L[m] == L[0].  The link of v (src/che.cpp:102-112) is L[0..m) for a closed
// fan and L[0..m] for an open one.  A vertex with EVT == NIL has no entries.
struct Che {
	uint32_t nv = 0, nf = 0, nhe = 0, ne = 0;
	bool manifold = true;
	std::vector<double> gt;     // 3*nv
	std::vector<uint32_t> vt, ot, evt;

	std::vector<uint32_t> fan_off;    // nv+1
	std::vector<uint32_t> fan;        // fan lists
	std::vector<uint8_t> fan_closed;  // 1 when the walk returned to EVT[v]
	double mean_edge = 1.0;           // mean length of the fan edges

	DeviceMesh *dev = nullptr;        // lazily uploaded device tables
	// Fan triangles taking the FP64 update (the flag test of build_layout) and
	// all fan slots of the closed-fan-without-repeat form; -1 until counted.
	mutable long long n_flag_tri = -1, n_fan_slots = -1;

	uint32_t n_tri(uint32_t v) const {
		const uint32_t c = fan_off[v + 1] - fan_off[v];
		return c ? c - 1 : 0;
	}
	uint32_t link_size(uint32_t v) const {
		const uint32_t c = fan_off[v + 1] - fan_off[v];
		return c ? (fan_closed[v] ? c - 1 : c) : 0;
	}
	~Che();
};

constexpr uint32_t ID_MASK = 0x7FFFFFFFu;
constexpr uint32_t F64_FLAG = 0x80000000u;

// FP32 geometry of fan entry k of a vertex, in units of the mean edge length h:
// x = |X_k|^2, y = X_k.X_{k+1} (y describes triangle (k, k+1) and is 0 on the
// last entry of an open fan; entry k + 1 is entry 0 after the last entry of a
// closed fan stored without its repeat).  The FP32 update forms the
// triangle's det and S from x, y and the next entry's x (update_stable_q).
struct Geo32 { float x, y; };
// FP64 geometry of the same entry: x, y as above, z = |X_k x X_{k+1}|^2,
// w = |X_{k+1} - X_k|^2 (FP64-fallback triangles).
struct Geo64 { double x, y, z, w; };
struct Edge64 { double x, y, z; };     // FP64 edge vector a - v of a fan entry

// FP64 update for triangle (k, k+1) of a fan when bit 31 of entry k is set:
// its angle at the updated vertex is below ~5.7 deg or above ~174 deg
// (sin^2 < 0.01: near-degenerate Gram matrix) or above ~166 deg (cos < -0.97,
// obtuse).  Measured on every corner of 512^2 and 1024^2 height fields with
// the FP64 distance field as inputs, the FP32 update of update_stable stays
// within 8e-7 of its FP64 evaluation (no corner of those meshes is flagged;
// the corners in 0.01 <= sin^2 < 0.03 or -0.97 <= cos < -0.9, which a wider
// test flags, are within 2.4e-7).  Also flagged: a corner whose two edges
// differ in length by more than 10x (squared lengths by more than
// FLAG_Q_RATIO), where the FP32 form in units of the mean edge loses
// precision -- on a 512^2 grid stretched 1:100 the unflagged corners' FP32
// updates were up to 8.6e-4 off their FP64 evaluation without this test and
// are within 3e-5 with it.  Regular meshes have no such corner; on a 1:100
// anisotropic grid 99% of the corners are flagged (89% by the angle tests).
constexpr double FLAG_SIN2_MIN = 0.01, FLAG_COS_MIN = -0.97, FLAG_Q_RATIO = 100.0;

// FP32 solves of a mesh whose flagged triangles exceed a quarter of its fan
// slots run FP64 sweeps instead (gp_ptp_opts.fp32_geometry == GP_GEO_AUTO):
// there most triangle updates already take the FP64 path, and an FP32 ring
// test can flip the schedule (a 1:100 stretched grid departs from the FP64
// schedule and ends 1.4e-3 off).  Counted once per mesh.
bool prefer_f64(const Che &m);

// ELL layout of a solve.  Positions (ring order: vertex order[pos], or pos
// when the mesh is numbered by rings) are grouped 64 to a chunk, one lane per
// vertex.  Chunk c has S_c = coff[c+1] - coff[c] slots, the longest fan in it;
// slot k of lane l lives at (coff[c] + k) * 64 + l, so every slot load of a
// wave is one contiguous 256-byte (id) or 512-byte (FP32 geometry) access.  Slot k holds fan entry
// k: cid = its vertex id, bit 31 = FP64 update for the triangle of slot k,
// bit 30 = empty slot (id = the vertex itself, never a triangle), bit 29 on
// slot 0 = the fan is closed and stored without its repeated first entry: the
// triangle of the last real slot joins it to slot 0 (layouts built with wrap).
// FP32 solves keep cgeo/cgeo64 (Geo32/Geo64 of entry k); FP64 solves keep
// cx64 = the FP64 edge vector a - v (24 bytes per slot).
constexpr uint32_t SLOT_EMPTY = 0x40000000u;
constexpr uint32_t SLOT_WRAP = 0x20000000u;    // also the vertex count limit of a layout
constexpr uint16_t LOCAL_OUT = 0xFFFF;
constexpr uint32_t SLOT_ID_MASK = 0x1FFFFFFFu;
struct Layout {
	uint32_t nchunks = 0, max_slots = 0;
	uint32_t uniform = 0;                  // every chunk has this many slots (0: see coff)
	unsigned long long n_flagged = 0;
	unsigned long long n_slots = 0;        // real slots of the positions
	std::vector<uint32_t> coff;
	std::vector<uint32_t> cid;
	std::vector<Geo32> cgeo;
	std::vector<Geo64> cgeo64;
	std::vector<Edge64> cx64;
	std::vector<unsigned long long> fcum;  // FP64-flagged triangles of positions [0, pos)
	std::vector<uint16_t> lidx;            // tile > 0: slot neighbour's index in the tile, or LOCAL_OUT
};
// tile > 0 also fills lidx (the mesh must be numbered by positions: order == null).
// W = 64: chunks of 64 positions, slot k of lane l at (coff[c] + k) * 64 + l
// (one wave per chunk, coalesced); W = 1: vertex-major, slot k of position p
// at coff[p] + k (one thread walks a vertex's slots in a few cache lines).
// uniform_max > 0 (W = 64 only): when no chunk needs more than uniform_max
// slots, every chunk gets the longest chunk's slot count, so chunk c starts at
// row c * max_slots and a kernel needs no coff read to find it (the extra
// slots are empty).  wrap: closed fans drop their repeated first entry (one
// slot less per interior vertex) and mark slot 0 with SLOT_WRAP; the kernels
// of relax_lane read both forms, the patch kernels only the plain one.
int build_layout(const Che &m, const uint32_t *order, uint32_t nr, bool f64, Layout &L, uint32_t tile = 0,
                 uint32_t W = 64, uint32_t uniform_max = 0, bool wrap = false);

// All vertices in Morton order of their positions, one scale for the three
// axes (a flat axis takes few bits), so consecutive vertices are close.
void morton_order(const Che &m, std::vector<uint32_t> &order);

// Band order for the band schedules: sources first (band 0), then reached
// vertices by band 1 + floor(dhat / width), inside a band by the Morton code
// of their position.  levels: toplesets (NIL = unreached).  Fills order
// (positions -> vertex) and limits (band starts, then the end).
int band_order(const Che &m, const uint32_t *src, uint32_t ns, const uint32_t *levels, const double *dhat,
               double width, std::vector<uint32_t> &order, std::vector<uint32_t> &limits);

// OFF / OBJ / PLY reader (mesh_io.cpp): positions [3*nv] and triangle corners.
int read_mesh(const char *path, std::vector<double> &xyz, std::vector<uint32_t> &faces);

// Builds VT/OT/EVT per the reference's rules and the fan tables.
int build_che(const double *xyz, uint32_t nv, const uint32_t *faces, uint32_t nf, Che &m);

// Host BFS rings (che::compute_toplesets).  limits is resized.
// k: the reference's level cut (NIL: none).
int toplesets_host(const Che &m, const uint32_t *src, uint32_t ns, uint32_t *levels,
                   uint32_t *sorted, std::vector<uint32_t> &limits, uint32_t k = NIL);

// Vertices renumbered by rings (ptp_coalescence, src/geodesics_ptp.cpp:15-38):
// the mesh of the reached vertices in ring order, plus inv (old -> new).
int coalesce(const Che &m, const uint32_t *sorted, uint32_t n_reached, Che &out,
             std::vector<uint32_t> &inv);

// FP64 host schedule (PTP_CPU semantics, last-read or last-written buffer).
int ptp_host(const Che &m, const uint32_t *src, uint32_t ns, const uint32_t *limits,
             uint32_t nl, const uint32_t *order, bool last_written, double tol, double *dist_out,
             uint32_t *sweeps_out);

// Device entry points (ptp_gpu.hip).
int device_check(char *name, size_t cap);
void device_mesh_free(DeviceMesh *d);
int toplesets_gpu(Che &m, const uint32_t *src, uint32_t ns, uint32_t *levels, uint32_t *sorted,
                  std::vector<uint32_t> &limits);
int plan_create_rings(Che *mesh, const uint32_t *src, uint32_t ns, const uint32_t *limits, uint32_t nl,
                      const uint32_t *sorted, const gp_ptp_opts &opts, gp_plan **out);
// Ring plan built on the device; with dhat, the rings are the nonempty bands
// of width `width` of dhat (the sources first), as the host band builder.
// With mlp (and no dhat), dhat is the network's prediction, computed on the
// device from the plan's own BFS levels (gp_plan_create_learned).
int plan_create_device(Che *mesh, const uint32_t *src, uint32_t ns, const gp_ptp_opts &opts, gp_plan **out,
                       const double *dhat = nullptr, double width = 0, gp_mlp *mlp = nullptr);
// BFS levels of the sources computed on the device (NIL where unreached).
int levels_gpu(Che &m, const uint32_t *src, uint32_t ns, uint32_t *levels);
int plan_run(gp_plan *p, gp_ptp_stats *stats);
// noalloc.hip: the allocation-free FP64 driver (stream: a hipStream_t).
int ptp_gpu_noalloc(Che &mesh, const uint32_t *d_src, uint32_t ns, const uint32_t *d_limits, uint32_t nl,
                    const uint32_t *d_sorted, double *d_dist, void *stream);
int plan_autotune(gp_plan *p, gp_tune *out);
// The band-width autotuner's remembered choice per mesh shape (vertex count
// rounded down to a power of two, longest link), schedule and precision.
bool tune_width_lookup(const Che &mesh, int schedule, int precision, double *width);
void tune_width_store(const Che &mesh, int schedule, int precision, double width);
int plan_result(gp_plan *p, double *dist);
int plan_clusters(gp_plan *p, uint32_t *clusters);
int ptp_iter_error(Che *mesh, const uint32_t *src, uint32_t ns, const double *exact, bool f64, uint32_t *iter,
                   double *err, uint32_t cap, uint32_t *n_err, uint32_t *band, uint32_t cap_band, uint32_t *n_band);
int plan_trace(gp_plan *p, uint32_t *out, uint32_t cap, uint32_t *n);

// Patch / band schedules (patches.hip).  dhat == null: no band gating.
struct PatchPlan;
int plan_create_patches(Che *mesh, const uint32_t *src, uint32_t ns, const double *dhat, const gp_ptp_opts &opts,
                        gp_plan **out);
int patch_create(Che *mesh, const uint32_t *src, uint32_t ns, const double *dhat, const gp_ptp_opts &o,
                 PatchPlan **out);
int patch_run(PatchPlan *pp, gp_ptp_stats *stats);
int patch_result(PatchPlan *pp, double *dist);
int patch_trace(PatchPlan *pp, uint32_t *out, uint32_t cap, uint32_t *n);
void patch_destroy(PatchPlan *pp);

const char *last_error();
void clear_error();

// Learned band predictor (mlp.hip).
constexpr uint32_t MLP_NPARAM = 64 * 32 + 64 + 64 * 64 + 64 + 64 + 1;
int fps_gpu(Che &m, uint32_t *samples, uint32_t n_init, uint32_t cap, uint64_t n, double radio, bool f64,
            uint32_t *n_out, double *radio_out, double *seconds);

int batch_create(Che *const *meshes, uint32_t nmeshes, const uint32_t *job_mesh, const uint32_t *job_source,
                 uint32_t njobs, double tol, uint32_t block, gp_batch **out);
int batch_run(gp_batch *b, gp_batch_stats *st);
int batch_result(const gp_batch *b, uint32_t job, double *dist, uint32_t cap, uint32_t *sweeps);

int mlp_create(uint32_t seed, gp_mlp **out);
int mlp_params(gp_mlp *m, float *get, const float *set, uint32_t n);
int mlp_features(Che &mesh, const uint32_t *src, uint32_t ns, const uint32_t *levels, float *F_out, float *e_out);
int mlp_forward(gp_mlp *m, const float *F, uint32_t n, float *y);
int mlp_train(gp_mlp *m, const float *F, const float *t, uint32_t n, uint32_t steps, uint32_t batch, float lr,
              uint32_t seed, float *grad_out, float *loss_out);
int mlp_predict(gp_mlp *m, Che &mesh, const uint32_t *src, uint32_t ns, const uint32_t *levels, double *dhat,
                float *ms_out);
int mlp_predict_device(gp_mlp *m, Che &mesh, const uint32_t *d_src, uint32_t ns, const uint32_t *d_levels,
                       uint32_t lmax, double *d_dhat, void *stream);

}  // namespace gp
