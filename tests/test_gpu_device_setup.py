"""Plans built on the device (gp_ptp_opts.device_setup) against plans built on
the host.

The device build replaces the host toplesets and renumbering by a frontier BFS
and a counting sort: the rings are the same sets, the order inside a ring is
not the reference's discovery order.  A sweep relaxes whole rings from the
previous buffer, so the schedule trace must be identical and the FP64 field
bit-identical to the host-built plan (and so to the FP64 oracle); FP32 fields
agree to the rounding of the mean edge length (the host build scales the
geometry by the mean edge of the renumbered mesh, summed in another order).
A fan longer than the register-held slots falls back to the host build.
"""
import numpy as np
import pytest

from conftest import rel_err
from gproshan_amd import _lib as L
from gproshan_amd import geodesics as P
from gproshan_amd import meshgen as G

pytestmark = pytest.mark.gpu

def _two_parts():
    """Two disjoint height fields, sources in the first only: the second part
    is never reached (+inf)."""
    x1, f1 = G.height_field(32, 5)
    x2, f2 = G.height_field(24, 6)
    return np.vstack([x1, x2 + np.array([3.0, 0.0, 0.0])]), np.vstack([f1, f2 + len(x1)])


CASES = [
    ("ico5", lambda: G.icosphere(5, 0.01, 0), [0]),
    ("hf128_4src", lambda: G.height_field(128, 1), [777, 5000, 9000, 16000]),
    ("hf256_16src", lambda: G.height_field(256, 7),
     G.draw_sources(np.random.default_rng(7), 256 * 256, 16).tolist()),
    ("aniso128", lambda: G.anisotropic_grid(128, 20.0, 0.4, 3), [4000]),
    ("polar10_rim", lambda: G.polar_disk(30, 10, 1), [1 + 29 * 10 + 3]),
    ("polar24", lambda: G.polar_disk(40, 24, 0), [0]),
    ("two_parts_one_reached", lambda: _two_parts(), [40]),
]


@pytest.fixture(scope="module", params=CASES, ids=[c[0] for c in CASES])
def case(request):
    name, gen, src = request.param
    xyz, f = gen()
    print(f"\n[case] {name} V={len(xyz)} sources={src}", flush=True)
    return name, np.array(src, np.uint32), P.che(xyz, f)


def _solve(C, src, **kw):
    plan = P.Plan(C, src, **kw)
    st = plan.run()
    d = plan.result()
    tr = plan.trace()
    st2 = plan.run()
    d2 = plan.result()
    plan.close()
    assert np.array_equal(d.view(np.uint64), d2.view(np.uint64)) and st2["sweeps"] == st["sweeps"]
    return d, tr, st


@pytest.mark.parametrize("sched", [dict(), dict(uncapped=True, tol=1e-6)], ids=["ref", "uncapped"])
def test_device_setup_f64_identical(case, oracle, sched):
    name, src, C = case
    host, tr_h, st_h = _solve(C, src, precision=L.PREC_F64, **sched)
    dev, tr_d, st_d = _solve(C, src, precision=L.PREC_F64, device_setup=True, **sched)
    print(f"[device f64] {name} {sched} sweeps {st_d['sweeps']}/{st_h['sweeps']}", flush=True)
    assert np.array_equal(host.view(np.uint64), dev.view(np.uint64))
    assert np.array_equal(tr_h, tr_d)


def test_device_setup_f32_close(case):
    name, src, C = case
    host, tr_h, _ = _solve(C, src, precision=L.PREC_F32)
    dev, tr_d, _ = _solve(C, src, precision=L.PREC_F32, device_setup=True)
    mx, _ = rel_err(dev, host)
    same = len(tr_h) == len(tr_d) and np.array_equal(tr_h[:, 1:7], tr_d[:, 1:7])
    print(f"[device f32] {name} max {mx:.2e} same schedule {same}", flush=True)
    assert mx <= (1e-6 if same else 1e-4)
    assert np.array_equal(np.isinf(host), np.isinf(dev))


def test_device_setup_errors():
    xyz, f = G.height_field(16, 0)
    C = P.che(xyz, f)
    with pytest.raises(P.GPError):
        P.Plan(C, [3, 3], device_setup=True)
    with pytest.raises(P.GPError):
        P.Plan(C, [len(xyz)], device_setup=True)


@pytest.mark.parametrize("kind", ["euclid", "dhat"])
def test_device_setup_bands_f64_identical(case, kind):
    """Band plans built on the device (bands of dhat as the rings, empty bands
    dropped) against the host band builder: same bands as sets, so the FP64
    field and the schedule trace are identical."""
    name, src, C = case
    if kind == "euclid":
        kw = dict(schedule=L.SCHED_EUCLID_BAND)
    else:
        rng = np.random.default_rng(5)
        xyz = C.gt()
        e = np.min(np.linalg.norm(xyz[:, None, :] - xyz[src][None, :, :], axis=2), axis=1)
        kw = dict(schedule=L.SCHED_LEARNED_BAND, dhat=e * rng.uniform(0.8, 1.3, len(e)))
    host, tr_h, _ = _solve(C, src, precision=L.PREC_F64, tol=1e-6, **kw)
    dev, tr_d, st = _solve(C, src, precision=L.PREC_F64, tol=1e-6, device_setup=True, **kw)
    print(f"[device bands f64] {name} {kind} sweeps {st['sweeps']} bands {len(tr_d)}", flush=True)
    assert np.array_equal(host.view(np.uint64), dev.view(np.uint64))
    assert np.array_equal(tr_h, tr_d)


def test_device_setup_bands_fall_back_on_large_band_index():
    xyz, f = G.height_field(32, 2)
    C = P.che(xyz, f)
    dhat = np.full(len(xyz), 1e300)
    host, tr_h, _ = _solve(C, [5], precision=L.PREC_F64, tol=1e-6, schedule=L.SCHED_LEARNED_BAND, dhat=dhat)
    dev, tr_d, _ = _solve(C, [5], precision=L.PREC_F64, tol=1e-6, schedule=L.SCHED_LEARNED_BAND, dhat=dhat,
                          device_setup=True)
    assert np.array_equal(host.view(np.uint64), dev.view(np.uint64))
    assert np.array_equal(tr_h, tr_d)
