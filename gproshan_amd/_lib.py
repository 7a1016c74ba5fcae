"""ctypes binding of libgproshan_ptp.so (the C ABI in include/gproshan_ptp.h).

The library is loaded from this package directory only.  If it is missing the
import fails with an explicit error: there is no fallback implementation.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgproshan_ptp.so")

NIL = 0xFFFFFFFF
PTP_TOL = 1e-3

OK, ERR_INVALID, ERR_NOMEM, ERR_DUPLICATE_SOURCE, ERR_HIP, ERR_NO_DEVICE, ERR_CAPACITY, ERR_KERNEL, \
    ERR_NOT_CONVERGED = 0, -1, -2, -3, -4, -5, -6, -7, -8
PREC_F32, PREC_F64 = 0, 1
RESULT_LAST_WRITTEN, RESULT_LAST_READ = 0, 1
SCHED_RINGS, SCHED_LEARNED_BAND, SCHED_EUCLID_BAND, SCHED_RING_BAND, SCHED_PATCHES = 0, 1, 2, 3, 4
LAUNCH_GRAPH, LAUNCH_STREAM, LAUNCH_FLOW = 0, 1, 2
GEO_AUTO, GEO_FP32, GEO_FP64 = 0, 1, 2

# Every symbol include/gproshan_ptp.h declares.
EXPORTS = [
    "gp_ptp_opts_default", "gp_last_error", "gp_device_check",
    "gp_che_create", "gp_che_destroy", "gp_che_read", "gp_che_positions", "gp_che_sizes", "gp_che_tables", "gp_che_link", "gp_che_mean_edge",
    "gp_compute_toplesets", "gp_compute_toplesets_k", "gp_compute_toplesets_gpu",
    "gp_parallel_toplesets_propagation_gpu", "gp_parallel_toplesets_propagation_coalescence_gpu",
    "gp_parallel_toplesets_propagation_gpu_noalloc",
    "gp_parallel_toplesets_propagation_cpu", "gp_parallel_toplesets_propagation_coalescence_cpu",
    "gp_normalize_ptp", "gp_ptp",
    "gp_plan_create", "gp_plan_create_bands", "gp_plan_run", "gp_plan_result", "gp_plan_trace", "gp_plan_destroy",
    "gp_plan_clusters", "gp_plan_autotune", "gp_plan_create_bands_tuned",
    "gp_mlp_create", "gp_mlp_destroy", "gp_mlp_params", "gp_mlp_features", "gp_mlp_forward", "gp_mlp_train",
    "gp_mlp_predict", "gp_plan_create_learned",
    "gp_batch_create", "gp_batch_run", "gp_batch_result", "gp_batch_destroy",
    "gp_farthest_point_sampling_ptp_gpu", "gp_ptp_iter_error",
]
MLP_NPARAM = 6337
MLP_NFEAT = 10


class GPError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"[{status}] {msg}")
        self.status = status


class PtpOpts(C.Structure):
    _fields_ = [("precision", C.c_int), ("result_buffer", C.c_int), ("schedule", C.c_int),
                ("coalesce", C.c_int), ("tol", C.c_double), ("max_sweeps", C.c_uint32),
                ("block", C.c_uint32), ("grid", C.c_uint32), ("launch_mode", C.c_int),
                ("band_width", C.c_double), ("local_iters", C.c_uint32), ("uncapped", C.c_int), ("clusters", C.c_int),
                ("device_setup", C.c_int), ("fp32_geometry", C.c_int), ("band_check", C.c_int)]


class PtpStats(C.Structure):
    _fields_ = [("sweeps", C.c_uint32), ("n_limits", C.c_uint32), ("relaxations", C.c_uint64),
                ("cond_checks", C.c_uint64), ("f64_updates", C.c_uint64), ("patch_visits", C.c_uint64),
                ("deferred", C.c_uint64), ("kernel_ms", C.c_double), ("total_ms", C.c_double),
                ("precision", C.c_int32), ("capped", C.c_int32), ("kernel", C.c_uint32), ("check_sweeps", C.c_uint32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


# Names of the sweep kernel variants (gp_ptp_stats.kernel bits: FP64, labels,
# eight register-held slots, FP64 geometry table), as rocprofv3 prints them.
def sweep_kernel_name(bits: int) -> str:
    f64, cl, ms8, g64 = bool(bits & 1), bool(bits & 2), bool(bits & 4), bool(bits & 8)
    b = lambda x: "true" if x else "false"
    if bits & 16:
        return f"gp::k_flow<{'8u' if ms8 else '10u'}, {b(g64)}>"
    tail = ", true" if g64 else ""
    return f"gp::k_sweep<{b(f64)}, {b(cl)}, {'8u' if ms8 else '10u'}{tail}>"


class Tune(C.Structure):
    _fields_ = [("grid", C.c_uint32), ("replay", C.c_uint32), ("kernel_ms", C.c_double),
                ("candidates", C.c_uint32), ("xcd_block", C.c_uint32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class BatchStats(C.Structure):
    _fields_ = [("sweeps", C.c_uint64), ("relaxations", C.c_uint64), ("vertex_fields", C.c_uint64),
                ("block", C.c_uint32), ("grid", C.c_uint32), ("kernel_ms", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`"
                          " (there is no fallback implementation)")
    L = C.CDLL(LIB_PATH)
    P, u32, i32 = C.c_void_p, C.c_uint32, C.c_int
    L.gp_ptp_opts_default.argtypes = [C.POINTER(PtpOpts)]
    L.gp_ptp_opts_default.restype = None
    L.gp_last_error.restype = C.c_char_p
    L.gp_device_check.argtypes = [C.c_char_p, C.c_size_t]
    L.gp_che_create.argtypes = [P, u32, P, u32, C.POINTER(P)]
    L.gp_che_read.argtypes = [C.c_char_p, C.POINTER(P)]
    L.gp_che_positions.argtypes = [P, P]
    L.gp_che_destroy.argtypes = [P]
    L.gp_che_destroy.restype = None
    L.gp_che_sizes.argtypes = [P, C.POINTER(u32), C.POINTER(u32), C.POINTER(u32), C.POINTER(u32), C.POINTER(i32)]
    L.gp_che_tables.argtypes = [P, P, P, P]
    L.gp_che_link.argtypes = [P, u32, P, u32, C.POINTER(u32)]
    L.gp_che_mean_edge.argtypes = [P, C.POINTER(C.c_double)]
    L.gp_compute_toplesets.argtypes = [P, P, u32, P, P, P, u32, C.POINTER(u32)]
    L.gp_compute_toplesets_gpu.argtypes = [P, P, u32, P, P, P, u32, C.POINTER(u32)]
    L.gp_compute_toplesets_k.argtypes = [P, P, u32, u32, P, P, P, u32, C.POINTER(u32)]
    drv = [P, P, u32, P, u32, P, P, P, C.POINTER(C.c_double)]
    L.gp_parallel_toplesets_propagation_gpu.argtypes = drv
    L.gp_parallel_toplesets_propagation_coalescence_gpu.argtypes = drv
    L.gp_parallel_toplesets_propagation_gpu_noalloc.argtypes = [P, P, u32, P, u32, P, P, P]
    L.gp_parallel_toplesets_propagation_cpu.argtypes = [P, P, u32, P, u32, P, P]
    L.gp_parallel_toplesets_propagation_coalescence_cpu.argtypes = [P, P, u32, P, u32, P, P]
    L.gp_normalize_ptp.argtypes = [P, u32]
    L.gp_ptp.argtypes = [P, P, u32, C.POINTER(PtpOpts), P, P, C.POINTER(PtpStats)]
    L.gp_plan_create.argtypes = [P, P, u32, C.POINTER(PtpOpts), C.POINTER(P)]
    L.gp_plan_create_bands.argtypes = [P, P, u32, P, C.POINTER(PtpOpts), C.POINTER(P)]
    L.gp_plan_run.argtypes = [P, C.POINTER(PtpStats)]
    L.gp_plan_result.argtypes = [P, P]
    L.gp_plan_trace.argtypes = [P, P, u32, C.POINTER(u32)]
    L.gp_plan_clusters.argtypes = [P, P]
    L.gp_plan_destroy.argtypes = [P]
    L.gp_plan_destroy.restype = None
    L.gp_plan_autotune.argtypes = [P, C.POINTER(Tune)]
    L.gp_plan_create_bands_tuned.argtypes = [P, P, u32, P, C.POINTER(PtpOpts), P, u32, C.POINTER(P),
                                             C.POINTER(C.c_double)]
    L.gp_mlp_create.argtypes = [u32, C.POINTER(P)]
    L.gp_mlp_destroy.argtypes = [P]
    L.gp_mlp_destroy.restype = None
    L.gp_mlp_params.argtypes = [P, P, P, u32]
    L.gp_mlp_features.argtypes = [P, P, u32, P, P]
    L.gp_mlp_forward.argtypes = [P, P, u32, P]
    L.gp_mlp_train.argtypes = [P, P, P, u32, u32, u32, C.c_float, u32, P, C.POINTER(C.c_float)]
    L.gp_mlp_predict.argtypes = [P, P, P, u32, P, C.POINTER(C.c_float)]
    L.gp_plan_create_learned.argtypes = [P, P, u32, P, C.POINTER(PtpOpts), C.POINTER(P)]
    L.gp_batch_create.argtypes = [P, u32, P, P, u32, C.c_double, u32, C.POINTER(P)]
    L.gp_batch_run.argtypes = [P, C.POINTER(BatchStats)]
    L.gp_batch_result.argtypes = [P, u32, P, u32, C.POINTER(u32)]
    L.gp_batch_destroy.argtypes = [P]
    L.gp_batch_destroy.restype = None
    L.gp_ptp_iter_error.argtypes = [P, P, u32, P, C.c_int, P, P, u32, C.POINTER(u32), P, u32, C.POINTER(u32)]
    L.gp_farthest_point_sampling_ptp_gpu.argtypes = [P, P, u32, u32, C.c_uint64, C.c_double, C.c_int,
                                                     C.POINTER(u32), C.POINTER(C.c_double), C.POINTER(C.c_double)]
    _lib = L
    return L


def check(rc: int):
    if rc != OK:
        raise GPError(rc, lib().gp_last_error().decode())


def ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def default_opts(**kw) -> PtpOpts:
    o = PtpOpts()
    lib().gp_ptp_opts_default(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o
