"""The dataflow launch (GP_LAUNCH_FLOW) against the sweep-per-launch path.

The dataflow launch runs the window schedule of src/geodesics_ptp.cpp:61-109
(and the uncapped loop of src/cuda/test_geodesics_ptp_coalescence.cu:223-248)
with the same windows, ring tests and relaxation as the graph-replayed sweeps;
only the synchronisation differs (each 64-vertex chunk waits for the
neighbour values it reads instead of a kernel boundary).  So the fields are
bit-identical to the graph path's (INF included), the sweep, relaxation,
ring-test and check-sweep counts and the per-sweep trace rows are equal, and
repeated runs are bit-identical (no dependence on timing).  The graph path is
itself pinned against the FP64 oracle elsewhere; one case here also checks the
dataflow field against the oracle's restatement of the reference schedule.
Plans the dataflow launch does not take (FP64 sweeps, cluster labels, the
sorted[] indirection of coalesce = 0) run the graph path and say so in
stats['kernel'] (bit 4 clear).
"""
import numpy as np
import pytest

from conftest import rel_err
from gproshan_amd import _lib as L
from gproshan_amd import geodesics as P
from gproshan_amd import meshgen as G
from gproshan_amd import workloads as W

pytestmark = pytest.mark.gpu

FLOW_BIT = 16
KEYS = ("sweeps", "relaxations", "cond_checks", "f64_updates", "check_sweeps", "capped", "precision")


def _run(mesh, src, dhat, launch, reps=1, **kw):
    plan = P.Plan(mesh, src, dhat, launch_mode=launch, **kw)
    out = []
    for _ in range(reps):
        st = plan.run()
        out.append((st, plan.result(), plan.trace()))
    plan.close()
    return out


def _check(mesh, src, dhat=None, expect_flow=True, reps=3, **kw):
    (sg, dg, tg), = _run(mesh, src, dhat, L.LAUNCH_GRAPH, **kw)
    runs = _run(mesh, src, dhat, L.LAUNCH_FLOW, reps=reps, **kw)
    assert not (sg["kernel"] & FLOW_BIT)
    for q, (sf, df, tf) in enumerate(runs):
        assert bool(sf["kernel"] & FLOW_BIT) == expect_flow, sf["kernel"]
        assert np.array_equal(df, dg, equal_nan=True), (q, float(np.nanmax(np.abs(df - dg))))
        for k in KEYS:
            assert sf[k] == sg[k], (q, k, sf[k], sg[k])
        assert np.array_equal(tf, tg), q
    sf = runs[0][0]
    print(f"  sweeps={sf['sweeps']} check={sf['check_sweeps']} graph {sg['kernel_ms']:.3f} ms, "
          f"flow {sf['kernel_ms']:.3f} ms", flush=True)
    return runs[0][1], sg


CASES = [
    ("ico5", lambda: G.icosphere(5, 0.01, 0), [0]),
    ("hf256", lambda: G.height_field(256, 2), [32896]),
    ("hf192_3src", lambda: G.height_field(192, 5), [100, 20000, 36000]),
    ("aniso128", lambda: G.anisotropic_grid(128, 20.0, 0.4, 3), [4000]),
]


@pytest.fixture(scope="module", params=CASES, ids=[c[0] for c in CASES])
def case(request):
    name, gen, src = request.param
    xyz, f = gen()
    print(f"\n[case] {name} V={len(xyz)} sources={src}", flush=True)
    return name, xyz, f, np.array(src, np.uint32), P.che(xyz, f)


@pytest.mark.parametrize("device_setup", [0, 1])
def test_flow_reference_schedule(case, device_setup):
    """The reference schedule (K/2 freeze, PTP_TOL, 2|limits| cap), FP32."""
    name, xyz, f, src, mesh = case
    _check(mesh, src, precision=L.PREC_F32, fp32_geometry=L.GEO_FP32, device_setup=device_setup)


def test_flow_uncapped(case):
    """The uncapped loop at tol 1e-6 (wide windows, many chunks per sweep)."""
    name, xyz, f, src, mesh = case
    _check(mesh, src, precision=L.PREC_F32, fp32_geometry=L.GEO_FP32, uncapped=True, tol=1e-6, device_setup=1)


def test_flow_euclid_bands(case):
    """Euclidean bands (0.5 mean edges) in place of the rings, with check sweeps."""
    name, xyz, f, src, mesh = case
    _check(mesh, src, None, precision=L.PREC_F32, fp32_geometry=L.GEO_FP32, schedule=L.SCHED_EUCLID_BAND,
           band_width=0.5, tol=1e-6, uncapped=True, device_setup=1)


def test_flow_fp64_geometry_table():
    """FP32 sweeps over the FP64 geometry table (every corner flagged)."""
    xyz, f = G.anisotropic_grid(96, 100.0, 0.4, 1)
    mesh = P.che(xyz, f)
    _check(mesh, np.array([4000], np.uint32), precision=L.PREC_F32, fp32_geometry=L.GEO_FP64)


def test_flow_not_taken_reports_graph():
    """FP64 sweeps, cluster labels and coalesce = 0 run the graph path."""
    xyz, f = G.height_field(128, 7)
    mesh = P.che(xyz, f)
    src = np.array([500], np.uint32)
    _check(mesh, src, expect_flow=False, reps=1, precision=L.PREC_F64)
    _check(mesh, src, expect_flow=False, reps=1, precision=L.PREC_F32, fp32_geometry=L.GEO_FP32, clusters=True)
    _check(mesh, src, expect_flow=False, reps=1, precision=L.PREC_F32, fp32_geometry=L.GEO_FP32, coalesce=False)


def test_flow_against_oracle(oracle):
    """The dataflow field of the reference schedule within 1e-4 of the FP64
    oracle's restatement of that schedule (same sweep count)."""
    xyz, f = G.height_field(256, 11)
    src = np.array([12345], np.uint32)
    ref, nsw = oracle.Mesh(xyz, f).ptp(src, oracle.GPU_COAL)
    mesh = P.che(xyz, f)
    d, st = _check(mesh, src, precision=L.PREC_F32, fp32_geometry=L.GEO_FP32, reps=1)
    mx, mean = rel_err(d, ref)
    print(f"  vs oracle: max {mx:.3e} mean {mean:.3e}", flush=True)
    assert mx <= 1e-4


@pytest.mark.parametrize("rows", ["plain_ref", "plain_uncapped", "euclid_band"])
def test_flow_full_size(rows):
    """The 1024^2 height field of the bench (1,048,576 vertices), one source."""
    xyz, f, src = W.c2(0, 1)
    mesh = P.che(xyz, f)
    kw = dict(precision=L.PREC_F32, fp32_geometry=L.GEO_FP32, device_setup=1)
    if rows == "plain_uncapped":
        kw.update(uncapped=True, tol=1e-6)
    if rows == "euclid_band":
        kw.update(uncapped=True, tol=1e-6, schedule=L.SCHED_EUCLID_BAND, band_width=0.5)
    print(f"\n[full size] {rows}", flush=True)
    _check(mesh, src, reps=2, **kw)


def test_flow_deadline_reports_error(monkeypatch):
    """A deadline shorter than the run makes the bounded waits give up: every
    wave leaves the kernel, the run returns GP_ERR_KERNEL, and the same plan
    runs normally (bit-identical to the graph path) once the deadline is back."""
    xyz, f = G.height_field(256, 3)
    mesh = P.che(xyz, f)
    src = np.array([20000], np.uint32)
    kw = dict(precision=L.PREC_F32, fp32_geometry=L.GEO_FP32, uncapped=True, tol=1e-6)
    (sg, dg, tg), = _run(mesh, src, None, L.LAUNCH_GRAPH, **kw)
    plan = P.Plan(mesh, src, None, launch_mode=L.LAUNCH_FLOW, **kw)
    monkeypatch.setenv("GPROSHAN_AMD_FLOW_DEADLINE_US", "2")
    with pytest.raises(P.GPError) as e:
        plan.run()
    assert e.value.status == L.ERR_KERNEL, e.value
    print(f"  short deadline: {e.value}", flush=True)
    monkeypatch.delenv("GPROSHAN_AMD_FLOW_DEADLINE_US")
    st = plan.run()
    assert st["kernel"] & FLOW_BIT
    assert np.array_equal(plan.result(), dg, equal_nan=True)
    assert st["sweeps"] == sg["sweeps"]
    plan.close()
