/*
 * gproshan_ptp.h -- C ABI of the MI355X (gfx950) parallel toplesets
 * propagation (PTP) geodesics library, libgproshan_ptp.so.
 *
 * Every entry point is extern "C", takes plain pointers and sizes, returns an
 * int status (GP_OK == 0, negative on error; gp_last_error() gives the text of
 * the last error on the calling thread) and never throws.  Each declaration
 * names the reference interface it replaces (path:line in larc/gproshan).
 *
 * Index type: uint32_t (index_t, include/include.h:19), GP_NIL = 0xFFFFFFFF
 * (NIL, include/include.h:12).  Distances cross the boundary as double
 * (distance_t = real_t = double in the reference's default build,
 * include/include.h:21-27); the arithmetic precision on the GPU is chosen by
 * gp_ptp_opts.precision.
 */
#ifndef GPROSHAN_PTP_H
#define GPROSHAN_PTP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GP_NIL 0xFFFFFFFFu
#define GP_PTP_TOL 1e-3  /* PTP_TOL, include/geodesics_ptp.h:11 */

enum gp_status {
	GP_OK = 0,
	GP_ERR_INVALID = -1,          /* bad argument (null pointer, index out of range, ...) */
	GP_ERR_NOMEM = -2,            /* host or device allocation failed                      */
	GP_ERR_DUPLICATE_SOURCE = -3, /* a source vertex is listed twice                       */
	GP_ERR_HIP = -4,              /* a HIP runtime call failed                             */
	GP_ERR_NO_DEVICE = -5,        /* no HIP device, or the device is not gfx950            */
	GP_ERR_CAPACITY = -6,         /* output buffer too small / size over the supported max */
	GP_ERR_KERNEL = -7,           /* a kernel reported an internal error (bounded wait hit) */
	GP_ERR_NOT_CONVERGED = -8     /* an uncapped or band loop stopped at its sweep guard
	                                 before every vertex met the tolerance              */
};

/* Arithmetic precision of the GPU relaxation. */
enum gp_precision {
	GP_PREC_F32 = 0,  /* FP32 bulk update; FP64 update on obtuse/near-degenerate triangles;
	                     FP64 convergence test.  The default of gp_ptp_opts_default.  With
	                     fp32_geometry == GP_GEO_AUTO, a mesh with more than a quarter of
	                     its fan triangles flagged runs FP64 sweeps instead (reported in
	                     gp_ptp_stats.precision).                                          */
	GP_PREC_F64 = 1   /* FP64 throughout, the reference's default build (include/config.h:7-8).
	                     The reference-named drivers (gp_parallel_toplesets_propagation_*_gpu)
	                     and the C++ front end default to it.                              */
};

/* Which double buffer the call returns.  The reference's CPU drivers return
 * the buffer the last sweep read (src/geodesics_ptp.cpp:113,193-198); its
 * GPU drivers return the buffer the last sweep wrote
 * (src/cuda/geodesics_ptp.cu:228-231, src/cuda/geodesics_ptp_coalescence.cu:156-159). */
enum gp_result_buffer {
	GP_RESULT_LAST_WRITTEN = 0,
	GP_RESULT_LAST_READ = 1
};

/* Schedule.  GP_SCHED_RINGS is the reference's window schedule over BFS
 * rings (gp_ptp_opts.uncapped selects its uncapped iter_error loop).  The
 * band schedules have no reference counterpart: a predicted distance dhat
 * groups the vertices into bands of band_width mean edge lengths (sources
 * first, Morton order inside a band), and the bands take the place of the
 * rings in the uncapped window loop -- the prediction only reorders the
 * propagation.  GP_SCHED_PATCHES runs the same FP32 relaxation to its fixed
 * point by Morton-ordered patches of `block` vertices iterated in LDS (up to
 * local_iters times per visit) with an active list of patches woken by
 * changed neighbours; given a dhat (gp_plan_create_bands) it only processes
 * the listed patches whose smallest prediction is within band_width mean edge
 * lengths of the list's smallest.  The patch schedule runs in FP32. */
enum gp_schedule {
	GP_SCHED_RINGS = 0,        /* toplesets (BFS rings), the reference schedule                     */
	GP_SCHED_LEARNED_BAND = 1, /* bands of the learned distance predictor (gp_mlp_*)                */
	GP_SCHED_EUCLID_BAND = 2,  /* bands of the Euclidean distance to the nearest source (ablation)  */
	GP_SCHED_RING_BAND = 3,    /* bands of BFS level * mean edge length (ablation: no predictor)    */
	GP_SCHED_PATCHES = 4       /* patch solver (optionally gated by a caller's dhat)                */
};

/* How the sweeps are launched.  GP_LAUNCH_FLOW runs the whole schedule of an
 * FP32 ring or band plan (no cluster labels, coalesce = 1) as one persistent
 * cooperative launch in which each 64-vertex chunk starts sweep k as soon as
 * the neighbour values of sweep k-1 it reads are published: the same windows,
 * ring tests and arithmetic, hence the same bits and sweep counts; other
 * plans run GP_LAUNCH_GRAPH (gp_ptp_stats.kernel tells which ran). */
enum gp_launch_mode {
	GP_LAUNCH_GRAPH = 0,   /* sweep + schedule kernels replayed from hipGraphs (default)          */
	GP_LAUNCH_STREAM = 1,  /* the same kernels launched one by one on the stream                 */
	GP_LAUNCH_FLOW = 2     /* one dataflow launch for the whole schedule                          */
};

/* Geometry table the FP32 ring sweeps read.  Both forms give the same bits:
 * the FP32 triangle updates use the FP32 rounding of the FP64 geometry and the
 * flagged corners (near-degenerate or obtuse, or with edges more than 10x
 * apart in length) the FP64 geometry itself.
 * GP_GEO_FP32 reads 8 bytes per slot (|X0|^2 and X0.X1; the FP32 update forms
 * the determinant and third edge from them) and loads the FP64 geometry of flagged
 * corners on the side (the first per vertex with the neighbour distances, the
 * others after the fan loop); GP_GEO_FP64 reads the 32-byte FP64 geometry of
 * every slot and runs each triangle once, which is faster when many corners
 * are flagged.  GP_GEO_AUTO picks GP_GEO_FP64 for host-built plans with more
 * than a quarter of their triangles flagged; since such meshes run FP64 sweeps
 * under GP_GEO_AUTO (gp_precision), GP_GEO_FP64 is only used when asked for. */
enum gp_fp32_geometry {
	GP_GEO_AUTO = 0,
	GP_GEO_FP32 = 1,
	GP_GEO_FP64 = 2
};

typedef struct gp_ptp_opts {
	int precision;      /* gp_precision                                          */
	int result_buffer;  /* gp_result_buffer                                      */
	int schedule;       /* gp_schedule                                           */
	int coalesce;       /* 1: renumber vertices by rings before the sweeps       */
	double tol;         /* relative-change tolerance (PTP_TOL for the ring test) */
	uint32_t max_sweeps;/* 0: the reference's 2*|limits| cap                     */
	uint32_t block;     /* patch schedules: patch size 256 or 512 (0: 256)       */
	uint32_t grid;      /* workgroups (0: autotuned/default)                     */
	int launch_mode;    /* gp_launch_mode                                        */
	double band_width;  /* band width / patch gate in mean edge lengths (0: 1 for
	                       bands, 128 for the patch gate)                       */
	uint32_t local_iters;/* patch schedules: iteration cap per patch visit       */
	int uncapped;       /* ring schedule: 1 = the reference's uncapped iter_error
	                       loop (src/cuda/test_geodesics_ptp_coalescence.cu:223-248):
	                       no K/2 freeze, no 2*|limits| cap; it runs to the
	                       relaxation's fixed point as tol -> 0               */
	int clusters;       /* ring schedule: 1 = also carry cluster (Voronoi) labels,
	                       the GPU drivers' rule (src/cuda/geodesics_ptp.cu:257-282):
	                       source q is q + 1, a vertex whose value drops takes the
	                       label of the nearer vertex of the winning triangle
	                       (old values), NIL where unreached                  */
	int device_setup;   /* ring schedule (coalesce = 1) and the learned and
	                       Euclidean band schedules, clusters = 0: 1 = build the
	                       rings (or bands) and the ring-numbered tables on the
	                       device (a frontier BFS and a counting sort instead of
	                       the host toplesets, band sort and renumbering; the
	                       order inside a ring is not the host's, which changes
	                       no value, and f64_updates is not counted); falls
	                       back to the host build when a fan is longer than the
	                       sweep kernels' register-held slots, a band index
	                       would pass the vertex count or the mesh is not
	                       manifold (the reference's coalescence rebuilds its
	                       half-edge mesh, which can change a star there)      */
	int fp32_geometry;  /* gp_fp32_geometry (precision F32, ring schedules)     */
	int band_check;     /* band schedules: 1 (default) = end every run with check
	                       sweeps over all reached vertices until none moves by
	                       tol (a vertex frozen in a band before its upwind
	                       neighbours settled is relaxed again); 0 = the bare
	                       band loop, as the uncapped loop over the same bands */
} gp_ptp_opts;

typedef struct gp_ptp_stats {
	uint32_t sweeps;          /* sweeps run (patch schedules: non-empty launches)     */
	uint32_t n_limits;        /* |limits| (patch schedules: number of patches)        */
	uint64_t relaxations;     /* vertex relaxations (window sizes; patch schedules:
	                             patch size x iterations per visit)                   */
	uint64_t cond_checks;     /* convergence checks (ring i sizes; patch visits x size) */
	uint64_t f64_updates;     /* FP64-fallback triangle updates (ring schedule, F32)  */
	uint64_t patch_visits;    /* patch schedules: patches processed                   */
	uint64_t deferred;        /* band schedules: patch entries carried over           */
	double kernel_ms;         /* device time of the sweep loop (hipEvent)             */
	double total_ms;          /* host wall time of the whole call                     */
	int32_t precision;        /* gp_precision the sweeps ran in                       */
	int32_t capped;           /* 1: the loop stopped at its sweep cap (the reference
	                             schedule's 2*|limits|, or the guard of the uncapped and
	                             band loops, which then return GP_ERR_NOT_CONVERGED)  */
	uint32_t kernel;          /* sweep kernel variant: bit 0 FP64, bit 1 cluster labels,
	                             bit 2 eight register-held slots, bit 3 FP64 geometry
	                             table (FP32 sweeps), bit 4 the dataflow launch ran  */
	uint32_t check_sweeps;    /* band schedules: full sweeps of the final convergence
	                             check over every reached vertex (1 when the bands
	                             already met tol)                                     */
} gp_ptp_stats;

/* Fill opts with the defaults (FP32, last written, rings, coalesce, PTP_TOL;
 * band_width 0 (schedule default), patches: block 256, local_iters 8,
 * band_check 1). */
void gp_ptp_opts_default(gp_ptp_opts *opts);

/* Text of the last error on this thread ("" if none). */
const char *gp_last_error(void);

/* Library/device probe: returns GP_OK if a gfx950 device is usable. */
int gp_device_check(char *name_out, size_t cap);

/* ---- compact half-edge mesh ------------------------------------------- */

typedef struct gp_che gp_che;

/* che::che(const vertex*, const index_t& n_v, const index_t* faces, const index_t& n_f)
 * -> che::init -> update_evt_ot_et (include/che.h:53; src/che.cpp:84-87,1254-1264,1295-1362).
 * xyz: n_vertices*3 doubles; faces: n_faces*3 vertex ids, consistently oriented. */
int gp_che_create(const double *xyz, uint32_t n_vertices, const uint32_t *faces,
                  uint32_t n_faces, gp_che **out);

/* che::~che (include/che.h:55; src/che.cpp:89-92,1404-1413). */
void gp_che_destroy(gp_che *mesh);

/* che_off / che_obj / che_ply (src/che_off.cpp:28-79, src/che_obj.cpp:24-79,
 * src/che_ply.cpp:25-156): reads a mesh file by extension (.off, .obj, .ply)
 * and builds it as gp_che_create does.  OFF quads split into (a,b,c),(d,a,c)
 * and OBJ quads into (a,b,d),(b,c,d), as the reference readers do; PLY is
 * ascii or binary_little_endian triangles.  Malformed files: GP_ERR_INVALID. */
int gp_che_read(const char *path, gp_che **out);

/* che::gt over all vertices (include/che.h:82): positions [3*n_vertices]. */
int gp_che_positions(const gp_che *mesh, double *xyz);

/* che::n_vertices/n_faces/n_half_edges/n_edges/is_manifold (include/che.h:80,93-96). */
int gp_che_sizes(const gp_che *mesh, uint32_t *n_vertices, uint32_t *n_faces,
                 uint32_t *n_half_edges, uint32_t *n_edges, int *manifold);

/* che::vt / che::ot / che::evt accessors over the whole tables
 * (include/che.h:81,89,91).  Any output may be NULL. */
int gp_che_tables(const gp_che *mesh, uint32_t *vt, uint32_t *ot, uint32_t *evt);

/* che::link (src/che.cpp:102-112): half-edge ids of the link of v. */
int gp_che_link(const gp_che *mesh, uint32_t v, uint32_t *he_out, uint32_t cap, uint32_t *n);

/* Mean edge length over the fans (no reference counterpart): the unit h of
 * the FP32 solve and of band_width. */
int gp_che_mean_edge(const gp_che *mesh, double *h);

/* ---- toplesets ----------------------------------------------------------- */

/* che::compute_toplesets(toplesets, sorted, limits, sources) (include/che.h:107;
 * src/che.cpp:546-593), on the host.  toplesets[n_vertices], sorted[n_vertices]
 * (unreached slots = GP_NIL), limits[cap_limits] with *n_limits = |limits|.
 * Duplicate sources: GP_ERR_DUPLICATE_SOURCE. */
int gp_compute_toplesets(const gp_che *mesh, const uint32_t *sources, uint32_t n_sources,
                         uint32_t *toplesets, uint32_t *sorted, uint32_t *limits,
                         uint32_t cap_limits, uint32_t *n_limits);

/* che::compute_toplesets with its level cut k (include/che.h:107 default
 * argument; src/che.cpp:572-575,592): the walk stops at the first vertex of
 * level k + 1, so sorted ends with the level-(k + 1) vertices discovered from
 * level k and the last ring [limits[n-2], limits[n-1]) holds levels k and
 * k + 1; toplesets is NIL beyond level k + 1.  k = GP_NIL: no cut. */
int gp_compute_toplesets_k(const gp_che *mesh, const uint32_t *sources, uint32_t n_sources, uint32_t k,
                           uint32_t *toplesets, uint32_t *sorted, uint32_t *limits,
                           uint32_t cap_limits, uint32_t *n_limits);

/* The same result computed on the GPU (frontier expansion + discovery-key
 * ordering); bit-identical to gp_compute_toplesets. */
int gp_compute_toplesets_gpu(const gp_che *mesh, const uint32_t *sources, uint32_t n_sources,
                             uint32_t *toplesets, uint32_t *sorted, uint32_t *limits,
                             uint32_t cap_limits, uint32_t *n_limits);

/* ---- PTP drivers (reference semantics) --------------------------------- */

/* parallel_toplesets_propagation_gpu(ptp_out, mesh, sources, toplesets)
 * (include/geodesics_ptp.h:36; src/cuda/geodesics_ptp.cu:20-85,174-299):
 * the window schedule over sorted[] on the original numbering; writes
 * dist[n_vertices] (INF where unreached) and returns the device seconds in
 * *seconds (may be NULL).  clusters (may be NULL) receives the cluster labels
 * (gp_ptp_opts.clusters). */
int gp_parallel_toplesets_propagation_gpu(gp_che *mesh, const uint32_t *sources, uint32_t n_sources,
                                          const uint32_t *limits, uint32_t n_limits,
                                          const uint32_t *sorted, double *dist,
                                          uint32_t *clusters, double *seconds);

/* parallel_toplesets_propagation_coalescence_gpu (include/geodesics_ptp.h:34;
 * src/cuda/geodesics_ptp_coalescence.cu:21-160): vertices renumbered by rings
 * (ptp_coalescence, src/geodesics_ptp.cpp:15-38) before the sweeps. */
int gp_parallel_toplesets_propagation_coalescence_gpu(gp_che *mesh, const uint32_t *sources,
                                                      uint32_t n_sources, const uint32_t *limits,
                                                      uint32_t n_limits, const uint32_t *sorted,
                                                      double *dist, uint32_t *clusters,
                                                      double *seconds);

/* No counterpart in the reference: the allocation-free form of
 * gp_parallel_toplesets_propagation_gpu, for callers that cannot allocate
 * device memory (code under HIP graph capture, serving loops that forbid
 * hipMalloc).  Same result (the FP64 window schedule over sorted[], K/2,
 * PTP_TOL, 2*|limits| cap, last-written buffer; bit-identical), from device
 * inputs: d_sources[n_sources], d_limits[n_limits] and d_sorted[n_vertices] as
 * gp_compute_toplesets writes them; the field goes to d_dist[n_vertices]
 * (device, INF where unreached).  Every kernel is enqueued on `stream` (a
 * hipStream_t; NULL: the null stream); the call allocates, frees and
 * synchronises nothing, so the caller waits on the stream.  The mesh's FP64
 * device tables must exist: the first FP64 solve with device_setup on the
 * mesh (gp_ptp or gp_plan_create with precision GP_PREC_F64, device_setup 1)
 * builds them; else GP_ERR_INVALID.  At most GP_NOALLOC_MAX_VERTICES vertices
 * (GP_ERR_CAPACITY beyond).  Its scratch memory (schedule state and two FP64
 * distance buffers, 16 MiB + 256 B) is one device array per GPU, held until
 * the process ends and shared by every call on that GPU: calls that overlap on
 * one GPU race. */
#define GP_NOALLOC_MAX_VERTICES 1048576u
int gp_parallel_toplesets_propagation_gpu_noalloc(gp_che *mesh, const uint32_t *d_sources, uint32_t n_sources,
                                                  const uint32_t *d_limits, uint32_t n_limits,
                                                  const uint32_t *d_sorted, double *d_dist, void *stream);

/* parallel_toplesets_propagation_cpu (include/geodesics_ptp.h:40;
 * src/geodesics_ptp.cpp:122-199): FP64 on the host (OpenMP), last-read buffer. */
int gp_parallel_toplesets_propagation_cpu(const gp_che *mesh, const uint32_t *sources,
                                          uint32_t n_sources, const uint32_t *limits,
                                          uint32_t n_limits, const uint32_t *sorted, double *dist);

/* parallel_toplesets_propagation_coalescence_cpu (include/geodesics_ptp.h:38;
 * src/geodesics_ptp.cpp:40-120): the PTP_CPU path of the front end. */
int gp_parallel_toplesets_propagation_coalescence_cpu(const gp_che *mesh, const uint32_t *sources,
                                                      uint32_t n_sources, const uint32_t *limits,
                                                      uint32_t n_limits, const uint32_t *sorted,
                                                      double *dist);

/* normalize_ptp (include/geodesics_ptp.h:46; src/geodesics_ptp.cpp:264-276). */
int gp_normalize_ptp(double *dist, uint32_t n);

/* ---- extended entry points (no single reference counterpart) ----------- */

/* geodesics(mesh, sources, PTP_GPU, dist) front-end dispatch
 * (src/geodesics.cpp:16-33,225-240) with explicit options: toplesets, then
 * the sweeps.  sorted_out[n_vertices] (may be NULL) receives the BFS order
 * the front end stores in sorted_index. */
int gp_ptp(gp_che *mesh, const uint32_t *sources, uint32_t n_sources, const gp_ptp_opts *opts,
           double *dist, uint32_t *sorted_out, gp_ptp_stats *stats);

/* Band-schedule plan from caller-supplied predicted distances dhat[n_vertices]
 * (opts->schedule must be a band schedule; the bands are 1 + floor(dhat /
 * (band_width * mean edge length)), sources form band 0). */
typedef struct gp_plan gp_plan;
int gp_plan_create_bands(gp_che *mesh, const uint32_t *sources, uint32_t n_sources, const double *dhat,
                         const gp_ptp_opts *opts, gp_plan **out);

/* Device-resident solver plan: mesh tables, bands and buffers live in HBM;
 * gp_plan_run() reruns the sweeps on them without host traffic (the timed
 * region of the benchmarks). */
int gp_plan_create(gp_che *mesh, const uint32_t *sources, uint32_t n_sources,
                   const gp_ptp_opts *opts, gp_plan **out);
int gp_plan_run(gp_plan *plan, gp_ptp_stats *stats);
int gp_plan_result(gp_plan *plan, double *dist);
int gp_plan_trace(gp_plan *plan, uint32_t *trace_out, uint32_t cap_sweeps, uint32_t *n_sweeps);
/* Cluster labels[n_vertices] of the last run (plans made with opts.clusters),
 * from the same buffer as gp_plan_result. */
int gp_plan_clusters(gp_plan *plan, uint32_t *clusters);
void gp_plan_destroy(gp_plan *plan);

/* ---- autotuner (no reference counterpart) ------------------------------ */

/* Launch shape of a ring or band plan's sweeps. */
typedef struct gp_tune {
	uint32_t grid;        /* workgroups of the sweep kernel (one wave each)            */
	uint32_t replay;      /* sweeps per replay of the long graph                       */
	double kernel_ms;     /* device time per field of the chosen shape                 */
	uint32_t candidates;  /* shapes timed (0: the choice came from the cache)          */
	uint32_t xcd_block;   /* log2 of the chunk blocks dealt round-robin to the eight
	                         XCDs (a chunk then stays on one XCD's L2 every sweep);
	                         0: each XCD takes a contiguous eighth of the window   */
} gp_tune;

/* Times the plan's sweep loop over candidate grids (a quarter, half, three
 * quarters and all of the workgroups that fit on the device at once) and
 * replay lengths (16, 32, 64, 128 sweeps), then the XCD deal of the best
 * (contiguous eighths, or blocks of 16, 64, 256 chunks): one warm-up run and
 * the mean device time of two runs each.  The fastest shape is kept in the plan and
 * remembered for plans of the same shape (vertex count rounded down to a power
 * of two, longest fan, precision, band or ring schedule, coalescing, uncapped
 * loop, schedule kind), which then take it without timing.  The plan's last result is that of
 * the last timed run.  Patch plans: GP_ERR_INVALID. */
int gp_plan_autotune(gp_plan *plan, gp_tune *out);

/* Band-width autotuner of the band schedules: builds the band plan of dhat
 * at each candidate width (mean edge lengths; widths == NULL: 0.45, 0.5, 0.6,
 * 0.75, 1.0), autotunes its launch shape (gp_plan_autotune), and returns the
 * plan with the smallest device time per field; *width_out (may be NULL)
 * receives its width.  Every run ends with the check sweeps (band_check), so
 * every candidate meets the same convergence test.  With the default
 * candidates the choice is remembered per mesh shape (vertex count rounded
 * down to a power of two, longest link), schedule and precision, and later
 * calls build that width's plan without timing.  Both autotuners keep their
 * choices for the process; when GPROSHAN_AMD_TUNE_FILE names a file, they are
 * also read from it and appended to it, so they carry over to later
 * processes. */
int gp_plan_create_bands_tuned(gp_che *mesh, const uint32_t *sources, uint32_t n_sources, const double *dhat,
                               const gp_ptp_opts *opts, const double *widths, uint32_t n_widths,
                               gp_plan **out, double *width_out);

/* ---- learned band predictor (no reference counterpart) ------------------ */

/* A 10 -> 64 -> 64 -> 1 ReLU MLP on per-vertex features, BF16 weights and
 * activations with FP32 accumulation (gfx950 MFMA), FP32 master weights and
 * Adam.  It predicts y = log(geodesic / Euclidean distance to the nearest
 * source); the band schedule orders vertices by dhat = e * exp(y). */
typedef struct gp_mlp gp_mlp;
#define GP_MLP_NPARAM 6337u  /* W1[64][32] (inputs padded to 32), b1[64], W2[64][64], b2[64], w3[64], b3 */
#define GP_MLP_NFEAT 10u

int gp_mlp_create(uint32_t seed, gp_mlp **out);
void gp_mlp_destroy(gp_mlp *mlp);
/* Copy the FP32 parameters out (get != NULL) and/or in (set != NULL); n == GP_MLP_NPARAM. */
int gp_mlp_params(gp_mlp *mlp, float *get, const float *set, uint32_t n);
/* Per-vertex features [n_vertices][GP_MLP_NFEAT] and Euclidean distance to the
 * nearest source, computed on the GPU. */
int gp_mlp_features(gp_che *mesh, const uint32_t *sources, uint32_t n_sources, float *features, float *euclid);
/* y[n] for feature rows F[n][GP_MLP_NFEAT]. */
int gp_mlp_forward(gp_mlp *mlp, const float *features, uint32_t n, float *y);
/* steps == 0: the gradient of mean((y - target)^2) over all n rows into
 * grad[GP_MLP_NPARAM] (no update).  steps > 0: that many Adam steps on
 * minibatches of `batch` rows (cosine-decayed lr).  *loss: last batch loss. */
int gp_mlp_train(gp_mlp *mlp, const float *features, const float *target, uint32_t n, uint32_t steps,
                 uint32_t batch, float lr, uint32_t seed, float *grad, float *loss);
/* dhat[n_vertices] for a mesh and sources (features + inference on the GPU);
 * *ms: device time of features + inference (may be NULL). */
int gp_mlp_predict(gp_mlp *mlp, gp_che *mesh, const uint32_t *sources, uint32_t n_sources, double *dhat,
                   float *ms);
/* Learned-band plan: gp_mlp_predict, then gp_plan_create_bands. */
int gp_plan_create_learned(gp_che *mesh, const uint32_t *sources, uint32_t n_sources, gp_mlp *mlp,
                           const gp_ptp_opts *opts, gp_plan **out);

/* iter_error_parallel_toplesets_propagation_coalescence_gpu
 * (src/cuda/test_geodesics_ptp_coalescence.cu:23-100, 214-266; used by
 * src/test_geodesics_ptp.cpp:196-213): the uncapped loop on the ring-numbered
 * mesh, one sweep at a time.  band[4 * cap_band] receives (sweep, i, j,
 * window size) for every sweep (the reference's "band" file); iter[cap] /
 * err[cap] receive (sweep, compute_error in %) after every sweep run once j is
 * the last ring (its "_error.iter" file).  exact[n_vertices] in the mesh's
 * numbering; every vertex must be reached.  *n_err / *n_band: the full counts
 * (entries beyond the capacities are not written). */
int gp_ptp_iter_error(gp_che *mesh, const uint32_t *sources, uint32_t n_sources, const double *exact,
                      int precision, uint32_t *iter, double *err, uint32_t cap, uint32_t *n_err,
                      uint32_t *band, uint32_t cap_band, uint32_t *n_band);

/* ---- farthest-point sampling ------------------------------------------- */

/* farthest_point_sampling_ptp_gpu (src/cuda/geodesics_ptp.cu:87-172,
 * include/geodesics_ptp.h:42): samples[cap] holds n_init distinct samples
 * (at least one) and receives the new ones; n is the total wanted (n >=
 * n_vertices means n_vertices / 2); rounds run while samples are wanted and
 * the last read maximum exceeds radio (read every round when radio > 0, else
 * in the last round), each appending the first vertex of largest distance of
 * the multi-source field of the current samples (plain GPU driver, PTP_TOL,
 * last-written buffer).  *n_out: samples now in the array; *radio_out: the
 * returned max distance (the reference's return value; INF if no round read
 * it); *seconds: wall time of the rounds.  precision GP_PREC_F64 reproduces
 * the reference bit for bit; GP_PREC_F32 is the FP32 solve with fallback.
 * Differences: n <= n_init adds nothing (the reference's size_t wraps), and a
 * round whose maximum is 0 stops instead of appending a duplicate. */
int gp_farthest_point_sampling_ptp_gpu(gp_che *mesh, uint32_t *samples, uint32_t n_init, uint32_t cap,
                                       uint64_t n, double radio, int precision, uint32_t *n_out,
                                       double *radio_out, double *seconds);

/* ---- batched single-source fields (no reference counterpart) ------------ */

/* Many independent one-source fields in one persistent launch: job q is the
 * field of job_source[q] on meshes[job_mesh[q]].  It replaces a host loop of
 * single-source calls, as in the reference's FPS and benchmark harness
 * (src/sampling.cpp:69, src/test_geodesics_ptp.cpp:310), each running the
 * window schedule of parallel_toplesets_propagation_gpu
 * (src/cuda/geodesics_ptp.cu:20-84) in FP32 with the FP64 fallback on
 * skinny/obtuse triangles, rings built on the GPU, last-written buffer.
 * block: threads per field (256, 512 or 1024), or 0 to autotune it by timing
 * the whole batch once per candidate. */
typedef struct gp_batch gp_batch;
typedef struct gp_batch_stats {
	uint64_t sweeps;         /* summed over the fields                         */
	uint64_t relaxations;
	uint64_t vertex_fields;  /* sum of the fields' vertex counts               */
	uint32_t block, grid;    /* launch shape in use                            */
	double kernel_ms;        /* device time of the launch                      */
} gp_batch_stats;

int gp_batch_create(gp_che *const *meshes, uint32_t n_meshes, const uint32_t *job_mesh,
                    const uint32_t *job_source, uint32_t n_jobs, double tol, uint32_t block, gp_batch **out);
int gp_batch_run(gp_batch *batch, gp_batch_stats *stats);
/* Field of job q (cap >= its mesh's vertex count); *sweeps may be NULL. */
int gp_batch_result(const gp_batch *batch, uint32_t job, double *dist, uint32_t cap, uint32_t *sweeps);
void gp_batch_destroy(gp_batch *batch);

#ifdef __cplusplus
}
#endif

#endif /* GPROSHAN_PTP_H */
