/*
 * TEST INFRASTRUCTURE -- NOT PART OF THE PRODUCT.
 *
 * CPU restatement of gproshan's parallel toplesets propagation (PTP) path,
 * written in plain C (C11 + OpenMP).  It is the checker for the HIP library:
 * only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py
 * load it (through ctypes).  The product library never links or calls it.
 *
 * Every function cites the reference file:line it restates.  Pins (what fixes
 * each function independently of this restatement) are listed per function;
 * the tests that exercise them live in tests/test_oracle_*.py.
 *
 * Floating point: the FP64 entry points follow the reference's default build
 * (real_t = double, include/include.h:21-25, include/config.h:7-8); the _f32
 * entry points follow the reference's SINGLE_P build (real_t = float).  The
 * oracle is compiled with -ffp-contract=off so that each product/sum is rounded
 * the way the reference's default x86-64 build rounds it (no FMA).
 */
#ifndef GP_ORACLE_H
#define GP_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_NIL 0xFFFFFFFFu

/* Compact half-edge mesh (reference: class che, include/che.h:27-130). */
typedef struct {
	uint32_t nv, nf, nhe, ne;
	double *gt;       /* 3*nv  geometry table            (include/che.h:41) */
	uint32_t *vt;     /* nhe   vertex table (faces)      (include/che.h:42) */
	uint32_t *ot;     /* nhe   opposite table            (include/che.h:43) */
	uint32_t *evt;    /* nv    extra vertex table        (include/che.h:44) */
	uint32_t *et;     /* ne    edge table                (include/che.h:45) */
	int manifold;     /* include/che.h:49                                   */
} or_che;

/* One sweep of the window schedule (the reference's "band" trace,
 * src/cuda/test_geodesics_ptp_coalescence.cu:245, extended with the window
 * bounds and the convergence count). */
typedef struct {
	uint32_t iter, i, j, start, end, ncond, count, pad;
} or_sweep;

/* Schedule variants. */
enum {
	OR_COALESCE   = 1,  /* renumber by rings + rebuild the mesh (src/geodesics_ptp.cpp:15-38) */
	OR_LAST_WRITE = 2,  /* return the buffer written by the last sweep (GPU drivers)          */
	OR_UNCAPPED   = 4   /* no K/2 band limit, no max_iter (iter_error trace loop)             */
};
/* Named combinations matching the reference drivers. */
#define OR_CPU_PLAIN  0                              /* parallel_toplesets_propagation_cpu             */
#define OR_CPU_COAL   (OR_COALESCE)                  /* parallel_toplesets_propagation_coalescence_cpu */
#define OR_GPU_PLAIN  (OR_LAST_WRITE)                /* parallel_toplesets_propagation_gpu             */
#define OR_GPU_COAL   (OR_COALESCE | OR_LAST_WRITE)  /* parallel_toplesets_propagation_coalescence_gpu */

/* Error codes (the reference has none; it asserts or reads out of bounds). */
#define OR_OK            0
#define OR_EINVAL       -1
#define OR_ENOMEM       -2
#define OR_EDUPSOURCE   -3
#define OR_ECAPACITY    -4

/* che::che(vertices, n_v, faces, n_f) -> init -> update_evt_ot_et
 * (src/che.cpp:84-87, 1254-1264, 1295-1362). xyz is FP64 [nv*3]. */
int  or_che_build(const double *xyz, uint32_t nv, const uint32_t *faces, uint32_t nf, or_che *out);
void or_che_free(or_che *m);

/* Half-edge index arithmetic (src/che.cpp:18-34). */
uint32_t or_next(uint32_t he);
uint32_t or_prev(uint32_t he);

/* che::star (src/che.cpp:94-100 with for_star, include/che.h:10) and
 * che::link (src/che.cpp:102-112).  Return the count; write at most cap. */
uint32_t or_star(const or_che *m, uint32_t v, uint32_t *out, uint32_t cap);
uint32_t or_link(const or_che *m, uint32_t v, uint32_t *out, uint32_t cap);

/* che::compute_toplesets (src/che.cpp:546-593).  levels[nv], sorted[nv],
 * limits[cap_limits]; k = NIL for no level cut.  Duplicate sources are
 * rejected with OR_EDUPSOURCE (the reference appends them twice and can write
 * past sorted[nv-1]). */
int or_toplesets(const or_che *m, const uint32_t *src, uint32_t ns, uint32_t k,
                 uint32_t *levels, uint32_t *sorted, uint32_t *limits,
                 uint32_t cap_limits, uint32_t *nlimits);

/* update_step (src/geodesics_ptp.cpp:201-262) on half-edge he. */
double or_update_step(const or_che *m, const double *dist, uint32_t he);
float  or_update_step_f32(const float *gt, const uint32_t *vt, const float *dist, uint32_t he);

/* The PTP window schedule (src/geodesics_ptp.cpp:40-199,
 * src/cuda/geodesics_ptp.cu:174-232, src/cuda/geodesics_ptp_coalescence.cu:102-160).
 * dist[nv] is fully written (INF where unreached).  tol is PTP_TOL (1e-3,
 * include/geodesics_ptp.h:11).  trace may be NULL.  n_sweeps receives the
 * number of sweeps run.  n_threads <= 0 means the OpenMP default. */
int or_ptp(const or_che *m, const uint32_t *src, uint32_t ns,
           const uint32_t *limits, uint32_t nlimits, const uint32_t *sorted,
           int variant, double tol, int n_threads,
           double *dist, or_sweep *trace, uint32_t trace_cap, uint32_t *n_sweeps);

/* The FP64 schedule with cluster labels (the GPU drivers' rule,
 * src/cuda/geodesics_ptp.cu:257-282): clusters[n_vertices] receives, from
 * the same buffer as the distances, the 1-based index of the source whose
 * front reached each vertex (NIL where unreached). */
int or_ptp_clusters(const or_che *m, const uint32_t *src, uint32_t ns,
                    const uint32_t *limits, uint32_t nlimits, const uint32_t *sorted,
                    int variant, double tol, int n_threads,
                    double *dist, uint32_t *clusters, uint32_t *n_sweeps);

/* The same schedule in the reference's SINGLE_P build (all float). */
int or_ptp_f32(const or_che *m, const uint32_t *src, uint32_t ns,
               const uint32_t *limits, uint32_t nlimits, const uint32_t *sorted,
               int variant, double tol, int n_threads,
               float *dist, or_sweep *trace, uint32_t trace_cap, uint32_t *n_sweeps);

/* Fixed point of the min-relaxation dist[v] = min(dist[v], min_star
 * update_step) that every un-truncated schedule converges to: a FIFO
 * label-correcting relaxation run until no value decreases by more than
 * rel_eps (relative).  Not a reference function: it is the limit of the
 * reference's algorithm when PTP_TOL -> 0 and no ring is frozen. */
int or_ptp_fixed_point(const or_che *m, const uint32_t *src, uint32_t ns,
                       double rel_eps, double *dist, uint64_t *n_relax);

/* farthest_point_sampling_ptp_gpu (src/cuda/geodesics_ptp.cu:87-172): the
 * sample budget n (>= n_vertices -> n_vertices / 2), then rounds of
 * compute_toplesets(samples) + the GPU plain driver (last-written buffer,
 * PTP_TOL) + the first index of the largest |distance| (cuBLAS I?amax), while
 * rounds remain and the last read maximum exceeds radio; the maximum is read
 * when radio > 0 or in the last round.  samples[cap] holds n_init distinct
 * samples and receives the new ones; a round whose maximum is 0 stops (the
 * reference would append a duplicate).  f32 != 0 runs the SINGLE_P arithmetic. */
int or_fps(const or_che *m, uint32_t *samples, uint32_t n_init, uint32_t cap, uint64_t n, double radio,
           int f32, uint32_t *n_out, double *radio_out);

/* iter_error_parallel_toplesets_propagation_coalescence_gpu
 * (src/cuda/test_geodesics_ptp_coalescence.cu:23-100, 214-266): the uncapped
 * loop on the ring-numbered mesh (every vertex must be reached); band[4*k] =
 * (sweep, i, j, end - start) for every sweep; (iter[k], err[k]) = (sweep,
 * compute_error in %) after every sweep run once j is the last ring.  The
 * counts are returned even when they exceed the capacities. */
int or_iter_error(const or_che *m, const uint32_t *src, uint32_t ns, const double *exact, int f32,
                  uint32_t *iter, double *err, uint32_t cap, uint32_t *n_err,
                  uint32_t *band, uint32_t cap_band, uint32_t *n_band);

/* compute_error (src/test_geodesics_ptp.cpp:362-371): mean relative error
 * in percent over n - s vertices. */
double or_compute_error(const double *dist, const double *exact, uint32_t n, uint32_t s);

/* normalize_ptp (src/geodesics_ptp.cpp:264-276). */
void or_normalize_ptp(double *dist, uint32_t n);

#ifdef __cplusplus
}
#endif

#endif
