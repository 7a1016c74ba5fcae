"""TEST INFRASTRUCTURE -- NOT PART OF THE PRODUCT.

ctypes binding of the C oracle (oracle/liboracle.so, built by oracle/Makefile).
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module.  See oracle/oracle.h for what each function restates.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

NIL = 0xFFFFFFFF
CPU_PLAIN, CPU_COAL, GPU_PLAIN, GPU_COAL = 0, 1, 2, 3
UNCAPPED = 4
PTP_TOL = 1e-3

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc only)."""
    if force and os.path.exists(_LIB):
        os.remove(_LIB)
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB


class _Che(C.Structure):
    _fields_ = [("nv", C.c_uint32), ("nf", C.c_uint32), ("nhe", C.c_uint32), ("ne", C.c_uint32),
                ("gt", C.POINTER(C.c_double)), ("vt", C.POINTER(C.c_uint32)),
                ("ot", C.POINTER(C.c_uint32)), ("evt", C.POINTER(C.c_uint32)),
                ("et", C.POINTER(C.c_uint32)), ("manifold", C.c_int)]


SWEEP_DTYPE = np.dtype([(k, np.uint32) for k in ("iter", "i", "j", "start", "end", "ncond", "count", "pad")])

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = C.CDLL(_LIB)
        P = C.c_void_p
        u32 = C.c_uint32
        L.or_che_build.argtypes = [P, u32, P, u32, C.POINTER(_Che)]
        L.or_che_free.argtypes = [C.POINTER(_Che)]
        L.or_star.argtypes = [C.POINTER(_Che), u32, P, u32]
        L.or_star.restype = u32
        L.or_link.argtypes = [C.POINTER(_Che), u32, P, u32]
        L.or_link.restype = u32
        L.or_toplesets.argtypes = [C.POINTER(_Che), P, u32, u32, P, P, P, u32, C.POINTER(u32)]
        L.or_update_step.argtypes = [C.POINTER(_Che), P, u32]
        L.or_update_step.restype = C.c_double
        L.or_update_step_f32.argtypes = [P, P, P, u32]
        L.or_update_step_f32.restype = C.c_float
        L.or_ptp.argtypes = [C.POINTER(_Che), P, u32, P, u32, P, C.c_int, C.c_double, C.c_int, P, P, u32,
                             C.POINTER(u32)]
        L.or_ptp_f32.argtypes = L.or_ptp.argtypes
        L.or_ptp_fixed_point.argtypes = [C.POINTER(_Che), P, u32, C.c_double, P, C.POINTER(C.c_uint64)]
        L.or_ptp_clusters.argtypes = [C.POINTER(_Che), P, u32, P, u32, P, C.c_int, C.c_double, C.c_int, P, P,
                                      C.POINTER(u32)]
        L.or_iter_error.argtypes = [C.POINTER(_Che), P, u32, P, C.c_int, P, P, u32, C.POINTER(u32), P, u32,
                                    C.POINTER(u32)]
        L.or_fps.argtypes = [C.POINTER(_Che), P, u32, u32, C.c_uint64, C.c_double, C.c_int, C.POINTER(u32),
                             C.POINTER(C.c_double)]
        L.or_compute_error.argtypes = [P, P, u32, u32]
        L.or_compute_error.restype = C.c_double
        L.or_normalize_ptp.argtypes = [P, u32]
        _lib = L
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


class Mesh:
    """Oracle compact half-edge mesh (che::che, src/che.cpp:84-87)."""

    def __init__(self, xyz, faces):
        self.xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        self.faces = np.ascontiguousarray(faces, dtype=np.uint32).reshape(-1)
        self._c = _Che()
        rc = lib().or_che_build(_p(self.xyz), len(self.xyz), _p(self.faces), len(self.faces) // 3,
                                C.byref(self._c))
        if rc != 0:
            raise ValueError(f"or_che_build failed: {rc}")
        nhe, nv = self._c.nhe, self._c.nv
        self.nv, self.nhe, self.ne = nv, nhe, self._c.ne
        self.vt = np.ctypeslib.as_array(self._c.vt, (max(nhe, 1),))[:nhe].copy()
        self.ot = np.ctypeslib.as_array(self._c.ot, (max(nhe, 1),))[:nhe].copy()
        self.evt = np.ctypeslib.as_array(self._c.evt, (nv,)).copy()
        self.manifold = bool(self._c.manifold)

    def __del__(self):
        try:
            lib().or_che_free(C.byref(self._c))
        except Exception:
            pass

    def star(self, v):
        out = np.zeros(256, np.uint32)
        n = lib().or_star(C.byref(self._c), v, _p(out), 256)
        return out[:n].copy()

    def link(self, v):
        out = np.zeros(256, np.uint32)
        n = lib().or_link(C.byref(self._c), v, _p(out), 256)
        return out[:n].copy()

    def toplesets(self, sources, k=NIL):
        src = np.ascontiguousarray(sources, dtype=np.uint32)
        levels = np.zeros(self.nv, np.uint32)
        sorted_ = np.zeros(self.nv, np.uint32)
        cap = self.nv + 2
        limits = np.zeros(cap, np.uint32)
        nl = C.c_uint32(0)
        rc = lib().or_toplesets(C.byref(self._c), _p(src), len(src), k, _p(levels), _p(sorted_),
                                _p(limits), cap, C.byref(nl))
        if rc != 0:
            raise ValueError(f"or_toplesets failed: {rc}")
        return levels, sorted_, limits[:nl.value].copy()

    def update_step(self, dist, he):
        d = np.ascontiguousarray(dist, dtype=np.float64)
        return lib().or_update_step(C.byref(self._c), _p(d), he)

    def ptp(self, sources, variant=CPU_PLAIN, tol=PTP_TOL, toplesets=None, f32=False, n_threads=0,
            trace=False):
        """Distances per the reference driver named by ``variant``.  Returns
        (dist, sweeps) or (dist, sweeps, trace records)."""
        src = np.ascontiguousarray(sources, dtype=np.uint32)
        if toplesets is None:
            toplesets = self.toplesets(src)
        _, sorted_, limits = toplesets
        limits = np.ascontiguousarray(limits, dtype=np.uint32)
        sorted_ = np.ascontiguousarray(sorted_, dtype=np.uint32)
        dist = np.zeros(self.nv, np.float32 if f32 else np.float64)
        cap = 2 * len(limits) + 8 if not (variant & UNCAPPED) else 64 * len(limits) + 64
        tr = np.zeros(cap, SWEEP_DTYPE) if trace else None
        nsw = C.c_uint32(0)
        fn = lib().or_ptp_f32 if f32 else lib().or_ptp
        rc = fn(C.byref(self._c), _p(src), len(src), _p(limits), len(limits), _p(sorted_), variant, tol,
                n_threads, _p(dist), _p(tr) if trace else None, cap if trace else 0, C.byref(nsw))
        if rc != 0:
            raise ValueError(f"or_ptp failed: {rc}")
        if trace:
            return dist, nsw.value, tr[:min(nsw.value, cap)].copy()
        return dist, nsw.value

    def ptp_clusters(self, sources, variant=GPU_COAL, tol=PTP_TOL):
        """(dist, clusters, sweeps) of the FP64 schedule with cluster labels."""
        src = np.ascontiguousarray(sources, dtype=np.uint32)
        _, sorted_, limits = self.toplesets(src)
        limits = np.ascontiguousarray(limits, dtype=np.uint32)
        sorted_ = np.ascontiguousarray(sorted_, dtype=np.uint32)
        dist = np.zeros(self.nv, np.float64)
        cl = np.zeros(self.nv, np.uint32)
        nsw = C.c_uint32(0)
        rc = lib().or_ptp_clusters(C.byref(self._c), _p(src), len(src), _p(limits), len(limits), _p(sorted_),
                                   variant, tol, 0, _p(dist), _p(cl), C.byref(nsw))
        if rc != 0:
            raise ValueError(f"or_ptp_clusters failed: {rc}")
        return dist, cl, nsw.value

    def iter_error(self, sources, exact, f32=False, cap=1 << 16):
        """((iter, err) pairs, band rows (sweep, i, j, window)) of the uncapped
        loop on the ring-numbered mesh."""
        src = np.ascontiguousarray(sources, dtype=np.uint32)
        ex = np.ascontiguousarray(exact, dtype=np.float64)
        it = np.zeros(cap, np.uint32)
        er = np.zeros(cap, np.float64)
        band = np.zeros((cap, 4), np.uint32)
        ne, nb = C.c_uint32(0), C.c_uint32(0)
        rc = lib().or_iter_error(C.byref(self._c), _p(src), len(src), _p(ex), int(f32), _p(it), _p(er), cap,
                                 C.byref(ne), _p(band), cap, C.byref(nb))
        if rc != 0:
            raise ValueError(f"or_iter_error failed: {rc}")
        return it[:min(ne.value, cap)], er[:min(ne.value, cap)], band[:min(nb.value, cap)]

    def fps(self, samples, n, radio=0.0, f32=False):
        """farthest_point_sampling_ptp_gpu: (samples, radio)."""
        init = np.ascontiguousarray(samples, dtype=np.uint32)
        cap = max(len(init), min(n, self.nv) + 1)
        buf = np.zeros(cap, np.uint32)
        buf[:len(init)] = init
        cnt = C.c_uint32(0)
        r = C.c_double(0)
        rc = lib().or_fps(C.byref(self._c), _p(buf), len(init), cap, n, radio, int(f32), C.byref(cnt), C.byref(r))
        if rc != 0:
            raise ValueError(f"or_fps failed: {rc}")
        return buf[:cnt.value].copy(), r.value

    def fixed_point(self, sources, rel_eps=1e-15):
        src = np.ascontiguousarray(sources, dtype=np.uint32)
        dist = np.zeros(self.nv, np.float64)
        nr = C.c_uint64(0)
        rc = lib().or_ptp_fixed_point(C.byref(self._c), _p(src), len(src), rel_eps, _p(dist), C.byref(nr))
        if rc != 0:
            raise ValueError(f"or_ptp_fixed_point failed: {rc}")
        return dist


def update_step_f32(gt, vt, dist, he):
    g = np.ascontiguousarray(gt, dtype=np.float32)
    v = np.ascontiguousarray(vt, dtype=np.uint32)
    d = np.ascontiguousarray(dist, dtype=np.float32)
    return lib().or_update_step_f32(_p(g), _p(v), _p(d), he)


def compute_error(dist, exact, s):
    d = np.ascontiguousarray(dist, dtype=np.float64)
    e = np.ascontiguousarray(exact, dtype=np.float64)
    return lib().or_compute_error(_p(d), _p(e), len(d), s)


def normalize_ptp(dist):
    d = np.ascontiguousarray(dist, dtype=np.float64).copy()
    lib().or_normalize_ptp(_p(d), len(d))
    return d
