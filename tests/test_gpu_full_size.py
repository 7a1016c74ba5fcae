"""The benchmark configurations at their full sizes, against the oracle.

  * C2: 1024 x 1024 height field (1,048,576 vertices), one source and 16
    sources: the reference window schedule in FP32 within 1e-4 of the FP64
    oracle of the same schedule (the one-source GPU path is the coalescence
    driver, several sources the sorted[] driver, src/geodesics.cpp:232-235),
    and FP64 bit-identical to it;
  * C5: 2048 x 2048 grid stretched 1:100 with jitter 0.4, 4 sources: an FP32
    request runs FP64 sweeps on this mesh (over a quarter of its triangles
    are flagged) and is bit-identical to the FP64 oracle;
  * C4: a held-out 1M-vertex height field: the learned-band schedule
    (0.6 mean edges wide, tol 1e-6, check sweeps) and the plain uncapped
    ring loop at the same tol, both within 1e-4 of the oracle's FP64 fixed
    point of the relaxation;
  * C3: 256 meshes of 2k-200k vertices, every (mesh, source) pair its own
    field, one launch: 64 fields drawn from the batch within 1e-4 of the
    FP64 oracle, every field reached where the oracle's is.
The oracle runs on the host cores (OpenMP).
"""
import os

import numpy as np
import pytest

from conftest import rel_err
from gproshan_amd import _lib as L
from gproshan_amd import geodesics as P
from gproshan_amd import workloads as W

pytestmark = pytest.mark.gpu

MODEL = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "gproshan_amd", "models",
                     "band_mlp.npy")


@pytest.mark.parametrize("n_sources", [1, 16])
def test_c2_reference_schedule(oracle, n_sources):
    xyz, f, src = W.c2(0, n_sources)
    print(f"\n[c2] V={len(xyz)} sources={src.tolist()}", flush=True)
    M = oracle.Mesh(xyz, f)
    ref, nsw = M.ptp(src, oracle.GPU_COAL if n_sources == 1 else oracle.GPU_PLAIN)
    C = P.che(xyz, f)
    for prec, bar in ((L.PREC_F32, 1e-4), (L.PREC_F64, 0.0)):
        plan = P.Plan(C, src, precision=prec, coalesce=n_sources == 1)
        st = plan.run()
        d = plan.result()
        plan.close()
        mx, mean = rel_err(d, ref)
        print(f"[c2] sources {n_sources} precision {prec}: sweeps {st['sweeps']}/{nsw} max {mx:.3e} mean {mean:.3e} "
              f"kernel {st['kernel_ms']:.2f} ms", flush=True)
        assert st["precision"] == prec
        assert mx <= bar
        if prec == L.PREC_F64:
            assert st["sweeps"] == nsw


def test_c5_stress(oracle):
    xyz, f, src = W.c5(0)
    print(f"\n[c5] V={len(xyz)} sources={src.tolist()}", flush=True)
    M = oracle.Mesh(xyz, f)
    ref, nsw = M.ptp(src, oracle.GPU_COAL)
    C = P.che(xyz, f)
    d, _, st = P.ptp(C, src, precision=L.PREC_F32)
    mx, _ = rel_err(d, ref)
    print(f"[c5] FP32 request ran precision {st['precision']}: sweeps {st['sweeps']}/{nsw} max {mx:.3e} "
          f"kernel {st['kernel_ms']:.2f} ms", flush=True)
    assert st["precision"] == L.PREC_F64
    assert st["sweeps"] == nsw
    assert mx == 0.0


@pytest.mark.skipif(not os.path.exists(MODEL), reason="no trained model")
def test_c4_learned_bands_held_out(oracle):
    from gproshan_amd import learned as Lr
    xyz, f, src = W.c4(0)
    print(f"\n[c4] V={len(xyz)} source={src.tolist()}", flush=True)
    fp = oracle.Mesh(xyz, f).fixed_point(src)
    C = P.che(xyz, f)
    net = Lr.MLP(params=np.load(MODEL, allow_pickle=False))
    dhat, _ = net.predict(C, src)
    for tag, kw in (("band_learned", dict(dhat=dhat, schedule=L.SCHED_LEARNED_BAND, band_width=0.6, tol=1e-6)),
                    ("plain_uncapped", dict(uncapped=True, tol=1e-6))):
        plan = P.Plan(C, src, **kw)
        st = plan.run()
        d = plan.result()
        plan.close()
        mx, mean = rel_err(d, fp)
        print(f"[c4] {tag}: sweeps {st['sweeps']} check sweeps {st['check_sweeps']} relaxations {st['relaxations']} "
              f"max {mx:.3e} mean {mean:.3e} kernel {st['kernel_ms']:.2f} ms", flush=True)
        assert st["capped"] == 0
        assert mx <= 1e-4


def test_c3_batch(oracle):
    raw, jobs = W.c3(0)
    meshes = [P.che(xyz, f) for xyz, f in raw]
    print(f"\n[c3] {len(meshes)} meshes, {len(jobs)} fields", flush=True)
    b = P.Batch(meshes, jobs, block=1024)
    st = b.run()
    print(f"[c3] kernel {st['kernel_ms']:.1f} ms sweeps {st['sweeps']} relaxations {st['relaxations']}", flush=True)
    pick = np.random.default_rng(3).choice(len(jobs), 64, replace=False)
    worst = 0.0
    for q in pick:
        m, s = jobs[q]
        ref, _ = oracle.Mesh(*raw[m]).ptp(np.array([s], np.uint32), oracle.GPU_PLAIN)
        d, _ = b.result(int(q))
        mx, _ = rel_err(d, ref)
        worst = max(worst, mx)
    b.close()
    print(f"[c3] 64 fields: max relative error vs the FP64 oracle {worst:.3e}", flush=True)
    assert worst <= 1e-4
