"""GPU cluster labels against the oracle's restatement of the reference GPU
drivers' rule (src/cuda/geodesics_ptp.cu:257-282): FP64 labels identical
(plain and coalesced drivers); FP32 labels equal on all
but a handful of vertices (a label can follow an FP32 tie the other way)."""
import numpy as np
import pytest

from conftest import rel_err
from gproshan_amd import _lib as L
from gproshan_amd import geodesics as P
from gproshan_amd import meshgen as G

pytestmark = pytest.mark.gpu

CASES = [
    ("hf128_4src", lambda: G.height_field(128, 1), [777, 5000, 9000, 16000]),
    ("ico5_3src", lambda: G.icosphere(5, 0.01, 0), [0, 4000, 9000]),
    ("hf256_16src", lambda: G.height_field(256, 7),
     G.draw_sources(np.random.default_rng(7), 256 * 256, 16).tolist()),
    ("polar24_3src", lambda: G.polar_disk(40, 24, 4), [0, 500, 900]),
]


@pytest.fixture(scope="module", params=CASES, ids=[c[0] for c in CASES])
def case(request, oracle):
    name, gen, src = request.param
    xyz, f = gen()
    return name, np.array(src, np.uint32), oracle.Mesh(xyz, f), P.che(xyz, f)


@pytest.mark.parametrize("coal", [True, False], ids=["coalesce", "plain"])
def test_clusters_f64_match_oracle(case, oracle, coal):
    name, src, M, C = case
    dref, cref, nsw = M.ptp_clusters(src, oracle.GPU_COAL if coal else oracle.GPU_PLAIN)
    plan = P.Plan(C, src, precision=L.PREC_F64, coalesce=coal, clusters=True)
    st = plan.run()
    d, cl = plan.result(), plan.clusters()
    mx, _ = rel_err(d, dref)
    print(f"\n[clusters f64] {name} coal={coal} sweeps {st['sweeps']}/{nsw} dist max {mx:.2e} "
          f"labels differ at {int((cl != cref).sum())}", flush=True)
    assert st["sweeps"] == nsw and mx == 0
    assert np.array_equal(cl, cref)


def test_clusters_f32_and_drivers(case, oracle):
    name, src, M, C = case
    dref, cref, _ = M.ptp_clusters(src, oracle.GPU_PLAIN)
    ts = C.compute_toplesets(src)
    d, secs, cl = P.parallel_toplesets_propagation_gpu(C, src, ts, clusters=True)
    diff = int((cl != cref).sum())
    mx, _ = rel_err(d, dref)
    print(f"[clusters f32] {name} dist max {mx:.2e} labels differ at {diff} of {len(cl)}", flush=True)
    assert mx <= 1e-4
    assert diff <= max(3, len(cl) // 10000)
    g = P.geodesics(C, src, P.geodesics.PTP_GPU, cluster=True)
    assert g.clusters is not None and len(g.clusters) == len(cl)
    assert int((g.clusters != cref).sum()) <= max(3, len(cl) // 10000)
