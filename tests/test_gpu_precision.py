"""Precision stress: a 1:100 stretched, jittered planar grid (obtuse and
skinny triangles everywhere), 4 sources, the reference schedule.

At 512 x 512 (the full 2048 x 2048 case is in tests/test_gpu_full_size.py):
  * all-FP64 GPU is bit-identical to the FP64 oracle;
  * an FP32 request runs FP64 sweeps on this mesh (over a quarter of its
    triangles are flagged, gp_precision) and is bit-identical too;
  * FP32 sweeps forced with the FP32 geometry table, the FP64 update on the
    flagged triangles, stay within 1e-4 of the FP64 oracle, and most of this
    mesh's triangles take the FP64 update;
  * the reference's own SINGLE_P arithmetic (all FP32, the oracle's f32
    build) is far off on the same mesh -- the failure the fallback exists for.
"""
import numpy as np
import pytest

from conftest import rel_err
from gproshan_amd import _lib as L
from gproshan_amd import geodesics as P
from gproshan_amd import meshgen as G

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def stress(oracle):
    xyz, f = G.anisotropic_grid(512, 100.0, 0.4, 5)
    src = G.draw_sources(np.random.default_rng(5), len(xyz), 4)
    M = oracle.Mesh(xyz, f)
    print(f"\n[stress] V={len(xyz)} sources={src.tolist()}", flush=True)
    return src, M, P.che(xyz, f)


def test_stress_f64_bit_exact(stress, oracle):
    src, M, C = stress
    ref, nsw = M.ptp(src, oracle.GPU_COAL)
    d, _, st = P.ptp(C, src, precision=L.PREC_F64)
    mx, _ = rel_err(d, ref)
    print(f"[stress f64] sweeps {st['sweeps']}/{nsw} max {mx:.2e} kernel {st['kernel_ms']:.2f} ms", flush=True)
    assert st["sweeps"] == nsw
    assert mx <= 1e-12


def test_stress_f32_request_runs_f64(stress, oracle):
    src, M, C = stress
    ref, nsw = M.ptp(src, oracle.GPU_COAL)
    d, _, st = P.ptp(C, src, precision=L.PREC_F32)
    mx, _ = rel_err(d, ref)
    print(f"[stress f32 request] ran precision {st['precision']} sweeps {st['sweeps']}/{nsw} max {mx:.2e}", flush=True)
    assert st["precision"] == L.PREC_F64
    assert st["sweeps"] == nsw
    assert mx <= 1e-12


def test_stress_f32_with_fallback(stress, oracle):
    src, M, C = stress
    ref, nsw = M.ptp(src, oracle.GPU_COAL)
    plan = P.Plan(C, src, precision=L.PREC_F32, fp32_geometry=L.GEO_FP32)
    st = plan.run()
    d = plan.result()
    plan.close()
    mx, mean = rel_err(d, ref)
    single, _ = M.ptp(src, oracle.GPU_COAL, f32=True)
    mxs, _ = rel_err(single.astype(np.float64), ref)
    frac = st["f64_updates"] / max(1, st["relaxations"] * 6)
    print(f"[stress f32] sweeps {st['sweeps']}/{nsw} max {mx:.2e} mean {mean:.2e}; FP64 share of triangle updates "
          f"~{frac:.2f}; all-FP32 reference arithmetic max {mxs:.2e}", flush=True)
    assert st["precision"] == L.PREC_F32 and st["f64_updates"] > 0
    assert mx <= 1e-4
    assert mxs > 10 * mx
