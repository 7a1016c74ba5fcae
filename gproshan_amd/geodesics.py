"""Python mirror of gproshan's PTP front end over the C ABI.

Names, argument meaning and error behaviour follow the reference:
  * ``che(vertices, faces)``                 -- include/che.h:53, src/che.cpp:84-87
  * ``che.compute_toplesets(sources)``       -- include/che.h:107, src/che.cpp:546-593
  * ``parallel_toplesets_propagation_*``     -- include/geodesics_ptp.h:34-40
  * ``normalize_ptp``                        -- include/geodesics_ptp.h:46
  * ``geodesics(mesh, sources, opt, ...)``   -- include/geodesics.h:18-72,
    src/geodesics.cpp:16-110,192-240 (PTP_CPU and PTP_GPU only)

Differences, by design: errors are raised as ``GPError`` instead of asserts or
out-of-bounds accesses (duplicate sources, an empty source list); the GPU
arithmetic precision and band schedule are keyword options; cluster labels
come from PTP_GPU only (the reference's CPU rule is a racy in-place update),
NIL where unreached.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import _lib as L
from ._lib import GPError, NIL, PTP_TOL, check, lib, ptr

__all__ = ["che", "geodesics", "parallel_toplesets_propagation_gpu",
           "parallel_toplesets_propagation_coalescence_gpu", "parallel_toplesets_propagation_cpu",
           "parallel_toplesets_propagation_coalescence_cpu", "normalize_ptp", "GPError", "NIL", "PTP_TOL",
           "Plan", "Batch", "farthest_point_sampling_ptp_gpu"]


def _u32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.uint32).reshape(-1))


class che:
    """Compact half-edge mesh (VT/OT/EVT tables built by the library)."""

    def __init__(self, vertices, faces):
        self._xyz = np.ascontiguousarray(np.asarray(vertices, dtype=np.float64).reshape(-1, 3))
        self._faces = _u32(faces)
        h = C.c_void_p()
        check(lib().gp_che_create(ptr(self._xyz), len(self._xyz), ptr(self._faces), len(self._faces) // 3,
                                  C.byref(h)))
        self._attach(h)

    @classmethod
    def read(cls, path):
        """che_off / che_obj / che_ply: a mesh file by extension (.off, .obj, .ply)."""
        h = C.c_void_p()
        check(lib().gp_che_read(str(path).encode(), C.byref(h)))
        obj = cls.__new__(cls)
        obj._attach(h)
        obj._xyz = np.zeros((obj._nv, 3))
        check(lib().gp_che_positions(h, ptr(obj._xyz)))
        obj._faces = obj.tables()[0].copy()
        return obj

    def _attach(self, h):
        self._h = h
        nv, nf, nhe, ne, man = C.c_uint32(), C.c_uint32(), C.c_uint32(), C.c_uint32(), C.c_int()
        check(lib().gp_che_sizes(self._h, C.byref(nv), C.byref(nf), C.byref(nhe), C.byref(ne), C.byref(man)))
        self._nv, self._nf, self._nhe, self._ne, self._manifold = nv.value, nf.value, nhe.value, ne.value, bool(man.value)
        self._tables = None

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().gp_che_destroy(h)
            self._h = None

    @property
    def handle(self):
        return self._h

    def n_vertices(self):
        return self._nv

    def n_faces(self):
        return self._nf

    def n_half_edges(self):
        return self._nhe

    def n_edges(self):
        return self._ne

    def is_manifold(self):
        return self._manifold

    def gt(self, v=None):
        return self._xyz if v is None else self._xyz[v]

    def tables(self):
        if self._tables is None:
            vt = np.zeros(max(self._nhe, 1), np.uint32)
            ot = np.zeros(max(self._nhe, 1), np.uint32)
            evt = np.zeros(self._nv, np.uint32)
            check(lib().gp_che_tables(self._h, ptr(vt), ptr(ot), ptr(evt)))
            self._tables = (vt[:self._nhe], ot[:self._nhe], evt)
        return self._tables

    def vt(self, he):
        return int(self.tables()[0][he])

    def ot(self, he):
        return int(self.tables()[1][he])

    def evt(self, v):
        return int(self.tables()[2][v])

    def link(self, v):
        out = np.zeros(256, np.uint32)
        n = C.c_uint32()
        check(lib().gp_che_link(self._h, v, ptr(out), 256, C.byref(n)))
        return out[:n.value].copy()

    def mean_edge(self):
        h = C.c_double()
        check(lib().gp_che_mean_edge(self._h, C.byref(h)))
        return h.value

    def compute_toplesets(self, sources, gpu: bool = False, k: int = NIL):
        """Returns (toplesets u32[V], sorted u32[V], limits u32[R+1]); k is the
        reference's level cut (host builder only)."""
        src = _u32(sources)
        top = np.zeros(self._nv, np.uint32)
        srt = np.zeros(self._nv, np.uint32)
        cap = self._nv + 2
        lim = np.zeros(cap, np.uint32)
        nl = C.c_uint32()
        if k != NIL:
            if gpu:
                raise GPError(L.ERR_INVALID, "the level cut k is computed by the host builder")
            check(lib().gp_compute_toplesets_k(self._h, ptr(src), len(src), int(k), ptr(top), ptr(srt), ptr(lim), cap,
                                               C.byref(nl)))
            return top, srt, lim[:nl.value].copy()
        fn = lib().gp_compute_toplesets_gpu if gpu else lib().gp_compute_toplesets
        check(fn(self._h, ptr(src), len(src), ptr(top), ptr(srt), ptr(lim), cap, C.byref(nl)))
        return top, srt, lim[:nl.value].copy()


def _driver(fn, mesh, sources, toplesets, gpu, clusters=False):
    src = _u32(sources)
    _, srt, lim = toplesets
    srt, lim = _u32(srt), _u32(lim)
    dist = np.full(mesh.n_vertices(), np.inf)
    if gpu:
        secs = C.c_double()
        cl = np.zeros(mesh.n_vertices(), np.uint32) if clusters else None
        check(fn(mesh.handle, ptr(src), len(src), ptr(lim), len(lim), ptr(srt), ptr(dist),
                 ptr(cl) if clusters else None, C.byref(secs)))
        return (dist, secs.value, cl) if clusters else (dist, secs.value)
    check(fn(mesh.handle, ptr(src), len(src), ptr(lim), len(lim), ptr(srt), ptr(dist)))
    return dist


def parallel_toplesets_propagation_gpu(mesh, sources, toplesets, clusters=False):
    """Returns (dist, seconds[, clusters]); sorted-indirected window sweeps on the GPU."""
    return _driver(lib().gp_parallel_toplesets_propagation_gpu, mesh, sources, toplesets, True, clusters)


def parallel_toplesets_propagation_coalescence_gpu(mesh, sources, toplesets, clusters=False):
    """Returns (dist, seconds[, clusters]); ring-renumbered window sweeps on the GPU."""
    return _driver(lib().gp_parallel_toplesets_propagation_coalescence_gpu, mesh, sources, toplesets, True,
                   clusters)


def parallel_toplesets_propagation_cpu(mesh, sources, toplesets):
    return _driver(lib().gp_parallel_toplesets_propagation_cpu, mesh, sources, toplesets, False)


def parallel_toplesets_propagation_coalescence_cpu(mesh, sources, toplesets):
    return _driver(lib().gp_parallel_toplesets_propagation_coalescence_cpu, mesh, sources, toplesets, False)


def farthest_point_sampling_ptp_gpu(mesh, samples, n, radio=0.0, precision=L.PREC_F64):
    """farthest_point_sampling_ptp_gpu (src/cuda/geodesics_ptp.cu:87-172):
    returns (samples, max distance, seconds); ``samples`` are the initial ones."""
    init = _u32(samples)
    nv = mesh.n_vertices()
    cap = max(len(init), min(int(n), nv) + 1)
    buf = np.zeros(cap, np.uint32)
    buf[:len(init)] = init
    cnt, r, secs = C.c_uint32(), C.c_double(), C.c_double()
    check(lib().gp_farthest_point_sampling_ptp_gpu(mesh.handle, ptr(buf), len(init), cap, int(n), float(radio),
                                                   int(precision), C.byref(cnt), C.byref(r), C.byref(secs)))
    return buf[:cnt.value].copy(), r.value, secs.value


def iter_error_parallel_toplesets_propagation_coalescence_gpu(mesh, sources, exact, precision=L.PREC_F64,
                                                              cap=1 << 16):
    """Returns (iter, err %, band rows (sweep, i, j, window)) of the uncapped
    loop (src/cuda/test_geodesics_ptp_coalescence.cu:23-100, 214-266)."""
    src = _u32(sources)
    ex = np.ascontiguousarray(exact, dtype=np.float64)
    it = np.zeros(cap, np.uint32)
    er = np.zeros(cap)
    band = np.zeros((cap, 4), np.uint32)
    ne, nb = C.c_uint32(), C.c_uint32()
    check(lib().gp_ptp_iter_error(mesh.handle, ptr(src), len(src), ptr(ex), int(precision), ptr(it), ptr(er), cap,
                                  C.byref(ne), ptr(band), cap, C.byref(nb)))
    return it[:min(ne.value, cap)], er[:min(ne.value, cap)], band[:min(nb.value, cap)]


def normalize_ptp(dist):
    d = np.ascontiguousarray(dist, dtype=np.float64)
    check(lib().gp_normalize_ptp(ptr(d), len(d)))
    return d


def ptp(mesh, sources, *, precision=L.PREC_F32, result_buffer=L.RESULT_LAST_WRITTEN, schedule=L.SCHED_RINGS,
        coalesce=True, tol=PTP_TOL, max_sweeps=0, grid=0, launch_mode=L.LAUNCH_GRAPH, uncapped=False,
        device_setup=False, fp32_geometry=L.GEO_AUTO):
    """One PTP solve with explicit options; returns (dist, sorted, stats dict).
    sorted is the reference's BFS order, built on the host while the solve runs."""
    src = _u32(sources)
    o = L.default_opts(precision=precision, result_buffer=result_buffer, schedule=schedule,
                       coalesce=int(coalesce), tol=tol, max_sweeps=max_sweeps, grid=grid,
                       launch_mode=launch_mode, uncapped=int(uncapped), device_setup=int(device_setup),
                       fp32_geometry=int(fp32_geometry))
    dist = np.zeros(mesh.n_vertices())
    srt = np.zeros(mesh.n_vertices(), np.uint32)
    st = L.PtpStats()
    check(lib().gp_ptp(mesh.handle, ptr(src), len(src), C.byref(o), ptr(dist), ptr(srt), C.byref(st)))
    return dist, srt, st.as_dict()


class Plan:
    """Device-resident solve (gp_plan_*): buffers stay in HBM between runs."""

    def __init__(self, mesh, sources, dhat=None, **kw):
        self.mesh = mesh
        self._src = _u32(sources)
        o = L.default_opts(**{k: (int(v) if k in ("coalesce", "uncapped", "clusters", "device_setup", "fp32_geometry",
                                                  "band_check") else v) for k, v in kw.items()})
        h = C.c_void_p()
        if dhat is None:
            check(lib().gp_plan_create(mesh.handle, ptr(self._src), len(self._src), C.byref(o), C.byref(h)))
        else:
            d = np.ascontiguousarray(dhat, dtype=np.float64)
            if d.shape != (mesh.n_vertices(),):
                raise GPError(L.ERR_INVALID, "dhat must hold one value per vertex")
            check(lib().gp_plan_create_bands(mesh.handle, ptr(self._src), len(self._src), ptr(d), C.byref(o),
                                             C.byref(h)))
        self._h = h

    @classmethod
    def tuned_bands(cls, mesh, sources, dhat, widths=None, **kw):
        """Band plan of dhat at the fastest of the candidate widths
        (gp_plan_create_bands_tuned); returns (plan, width)."""
        obj = cls.__new__(cls)
        obj.mesh = mesh
        obj._src = _u32(sources)
        o = L.default_opts(**{k: (int(v) if k in ("coalesce", "uncapped", "clusters", "device_setup", "fp32_geometry",
                                                  "band_check") else v) for k, v in kw.items()})
        d = np.ascontiguousarray(dhat, dtype=np.float64)
        if d.shape != (mesh.n_vertices(),):
            raise GPError(L.ERR_INVALID, "dhat must hold one value per vertex")
        w = None if widths is None else np.ascontiguousarray(widths, dtype=np.float64)
        h, wout = C.c_void_p(), C.c_double()
        check(lib().gp_plan_create_bands_tuned(mesh.handle, ptr(obj._src), len(obj._src), ptr(d), C.byref(o),
                                               None if w is None else ptr(w), 0 if w is None else len(w),
                                               C.byref(h), C.byref(wout)))
        obj._h = h
        return obj, wout.value

    @classmethod
    def learned(cls, mesh, sources, mlp, **kw):
        """Learned-band plan in one call (gp_plan_create_learned): the network's
        prediction (features and forward on the device) ordered into bands."""
        obj = cls.__new__(cls)
        obj.mesh = mesh
        obj._src = _u32(sources)
        o = L.default_opts(**{k: (int(v) if k in ("coalesce", "uncapped", "clusters", "device_setup", "fp32_geometry",
                                                  "band_check") else v) for k, v in kw.items()})
        h = C.c_void_p()
        check(lib().gp_plan_create_learned(mesh.handle, ptr(obj._src), len(obj._src), mlp.handle, C.byref(o),
                                           C.byref(h)))
        obj._h = h
        return obj

    def autotune(self):
        """gp_plan_autotune: the plan's launch shape (grid, replay) timed and kept."""
        t = L.Tune()
        check(lib().gp_plan_autotune(self._h, C.byref(t)))
        return t.as_dict()

    def run(self):
        st = L.PtpStats()
        check(lib().gp_plan_run(self._h, C.byref(st)))
        return st.as_dict()

    def result(self):
        d = np.zeros(self.mesh.n_vertices())
        check(lib().gp_plan_result(self._h, ptr(d)))
        return d

    def clusters(self):
        c = np.zeros(self.mesh.n_vertices(), np.uint32)
        check(lib().gp_plan_clusters(self._h, ptr(c)))
        return c

    def trace(self):
        n = C.c_uint32()
        check(lib().gp_plan_trace(self._h, None, 0, C.byref(n)))
        out = np.zeros((max(n.value, 1), 8), np.uint32)
        check(lib().gp_plan_trace(self._h, ptr(out), n.value, C.byref(n)))
        return out[:n.value]

    def close(self):
        if getattr(self, "_h", None):
            lib().gp_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


class Batch:
    """Many single-source fields in one persistent launch (gp_batch_*).

    jobs: sequence of (mesh index, source vertex).  block: threads per field
    (256, 512, 1024) or 0 to autotune it on the batch itself.
    """

    def __init__(self, meshes, jobs, tol=L.PTP_TOL, block=0):
        self.meshes = list(meshes)
        jobs = np.asarray(jobs, dtype=np.int64).reshape(-1, 2)
        self._jm = _u32(jobs[:, 0])
        self._js = _u32(jobs[:, 1])
        hs = (C.c_void_p * max(len(self.meshes), 1))(*[m.handle for m in self.meshes])
        h = C.c_void_p()
        check(lib().gp_batch_create(hs, len(self.meshes), ptr(self._jm), ptr(self._js), len(jobs), float(tol),
                                    int(block), C.byref(h)))
        self._h = h

    def run(self):
        st = L.BatchStats()
        check(lib().gp_batch_run(self._h, C.byref(st)))
        return st.as_dict()

    def result(self, job):
        n = self.meshes[int(self._jm[job])].n_vertices()
        d = np.zeros(n)
        sw = C.c_uint32()
        check(lib().gp_batch_result(self._h, int(job), ptr(d), n, C.byref(sw)))
        return d, sw.value

    def close(self):
        if getattr(self, "_h", None):
            lib().gp_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


class geodesics:
    """gproshan::geodesics restricted to the PTP options.

    The option numbering is the reference's CUDA build (GPROSHAN_CUDA defined,
    include/geodesics.h:21-28).  FM, HEAT_FLOW and HEAT_FLOW_GPU are other
    algorithms, outside this library: requesting them raises.  PTP_GPU computes
    in FP64 by default, as the reference's default build (include/config.h:7-8);
    precision=PREC_F32 selects the FP32 bulk update with the FP64 fallback.
    """
    FM, PTP_GPU, HEAT_FLOW_GPU, PTP_CPU, HEAT_FLOW = range(5)

    def __init__(self, mesh, sources, opt=FM, e_dist=None, cluster=False, n_iter=0, radio=math.inf, *,
                 precision=L.PREC_F64, schedule=L.SCHED_RINGS):
        n = mesh.n_vertices()
        if n <= 0:
            raise GPError(L.ERR_INVALID, "mesh has no vertices")
        src = _u32(sources)
        if len(src) == 0:
            raise GPError(L.ERR_INVALID, "no sources")            # src/geodesics.cpp:31
        if cluster and opt != self.PTP_GPU:
            raise GPError(L.ERR_INVALID, "cluster labels are produced by PTP_GPU only")
        if e_dist is not None and (not isinstance(e_dist, np.ndarray) or e_dist.dtype != np.float64
                                   or e_dist.size < n or not e_dist.flags.c_contiguous):
            raise GPError(L.ERR_INVALID, "e_dist must be a contiguous float64 array of n_vertices")
        self.clusters = None
        self._dist = e_dist if e_dist is not None else np.empty(n)
        self._dist[:n] = np.inf                                     # src/geodesics.cpp:28-29
        self._sorted = np.full(n, NIL, np.uint32)                   # :27
        self._n_sorted = 0
        self._n = n
        if opt == self.PTP_CPU:
            ts = mesh.compute_toplesets(src)                        # src/geodesics.cpp:192-207
            self._sorted[:] = ts[1]
            self._dist[:n] = parallel_toplesets_propagation_coalescence_cpu(mesh, src, ts)
        elif opt == self.PTP_GPU and cluster:
            # The reference's GPU drivers with their cluster output
            # (src/geodesics.cpp:225-240).
            ts = mesh.compute_toplesets(src)
            drv = parallel_toplesets_propagation_coalescence_gpu if len(src) == 1 else parallel_toplesets_propagation_gpu
            d, _, cl = drv(mesh, src, ts, clusters=True)
            self._sorted[:] = ts[1]
            self._dist[:n] = d
            self.clusters = cl
        elif opt == self.PTP_GPU:
            # The reference runs the coalescence driver for one source and the
            # sorted[] driver for several (src/geodesics.cpp:225-240); both give
            # the same field.  Here both run ring-numbered sweeps whose rings
            # and tables are built on the device (the host builds them for
            # other schedules and for fans past the register-held slots).
            d, srt, _ = ptp(mesh, src, precision=precision, schedule=schedule, coalesce=True,
                            device_setup=schedule == L.SCHED_RINGS)
            self._sorted[:] = srt
            self._dist[:n] = d
        else:
            raise GPError(L.ERR_INVALID, f"option {opt} is not a PTP option")

    def __getitem__(self, i):
        if not 0 <= i < self._n:
            raise IndexError(i)
        return self._dist[i]

    def __call__(self, i):
        if not 0 <= i < self._n:
            raise IndexError(i)
        return int(self._sorted[i])

    def n_sorted_index(self):
        return self._n_sorted

    def copy_sorted_index(self, n):
        if n > self._n_sorted:
            raise GPError(L.ERR_INVALID, "n exceeds n_sorted_index()")
        return self._sorted[:n].copy()

    def farthest(self):
        if self._n_sorted == 0:                                     # src/geodesics.cpp:60-64 asserts
            raise GPError(L.ERR_INVALID, "farthest() needs a sorted run (FM); PTP leaves n_sorted = 0")
        return int(self._sorted[self._n_sorted - 1])

    def radio(self):
        return self._dist[self.farthest()]

    def normalize(self):
        # n_sorted is 0 after PTP, so the reference falls back to normalize_ptp
        # (src/geodesics.cpp:77-90).
        self._dist[:self._n] = normalize_ptp(self._dist[:self._n])

    @property
    def dist(self):
        return self._dist[:self._n]
