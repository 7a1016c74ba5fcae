"""GPU band and patch schedules.

Band schedules: a predicted distance groups the vertices into bands that take
the place of the BFS rings in the reference's uncapped window loop
(src/cuda/test_geodesics_ptp_coalescence.cu:223-248).  In FP64 the GPU result
is bit-identical to the oracle running that loop over the same band order
(rebuilt here from the documented rule: sources first, then band
floor(dhat / width), then the 30-bit Morton code of the position on the
per-axis normalised 1024^3 lattice, then the vertex id).

Against the oracle's FP64 fixed point of the relaxation: the patch schedule
at tol 1e-6 lands within 1e-4 whatever the prediction (exact, useless or
adversarial) -- a prediction only reorders its work.  The band schedules are
window loops: a band retires after one sweep without a change above tol, so a
vertex whose upwind neighbours were predicted into a later band can be frozen
early (up to 1.4e-4 off on a 256^2 height field with 1-edge Euclidean bands
without the final check); the check sweeps over every vertex that end a band
run continue until no vertex moves by tol, so at tol 1e-6 every band
schedule lands within 1e-4.
"""
BAND_BAR = 1e-4
import numpy as np
import pytest

from conftest import rel_err
from gproshan_amd import _lib as L
from gproshan_amd import geodesics as P
from gproshan_amd import meshgen as G

pytestmark = pytest.mark.gpu

CASES = [
    ("ico5", lambda: G.icosphere(5, 0.01, 0), [0]),
    ("hf256", lambda: G.height_field(256, 2), [32896]),
    ("hf192_3src", lambda: G.height_field(192, 5), [100, 20000, 36000]),
    ("aniso128", lambda: G.anisotropic_grid(128, 20.0, 0.4, 3), [4000]),
]


@pytest.fixture(scope="module", params=CASES, ids=[c[0] for c in CASES])
def case(request, oracle):
    name, gen, src = request.param
    xyz, f = gen()
    print(f"\n[case] {name} V={len(xyz)} sources={src}", flush=True)
    M = oracle.Mesh(xyz, f)
    return name, xyz, f, np.array(src, np.uint32), M, P.che(xyz, f), M.fixed_point(src)


def _spread10(x):
    x = x.astype(np.uint64) & 0x3FF
    x = (x | (x << 16)) & 0x030000FF
    x = (x | (x << 8)) & 0x0300F00F
    x = (x | (x << 4)) & 0x030C30C3
    x = (x | (x << 2)) & 0x09249249
    return x


def band_order(xyz, src, levels, dhat, width):
    lo, hi = xyz.min(0), xyz.max(0)
    sc = np.where(hi > lo, 1023.0 / np.where(hi > lo, hi - lo, 1.0), 0.0)
    q = ((xyz - lo) * sc).astype(np.uint32)
    mort = _spread10(q[:, 0]) | (_spread10(q[:, 1]) << 1) | (_spread10(q[:, 2]) << 2)
    is_src = np.zeros(len(xyz), bool)
    is_src[src] = True
    keep = np.flatnonzero((levels != 0xFFFFFFFF) & ~is_src)
    d = dhat[keep]
    bq = np.where(np.isfinite(d) & (d > 0), np.floor(np.where(np.isfinite(d), d, 0) / width), 0.0)
    band = 1 + np.minimum(bq, 4.0e9).astype(np.uint64)
    key = (band << 30) | mort[keep]
    o = keep[np.lexsort((keep, key))]
    order = np.concatenate([src, o]).astype(np.uint32)
    kb = np.concatenate([np.zeros(len(src), np.uint64), np.sort(key) >> 30])
    limits = np.concatenate([[0], np.flatnonzero(kb[1:] != kb[:-1]) + 1, [len(order)]]).astype(np.uint32)
    return order, limits


@pytest.mark.parametrize("width", [1.0, 4.0])
def test_euclid_bands_f64_match_oracle(case, oracle, width):
    name, xyz, f, src, M, C, fp = case
    levels, _, _ = M.toplesets(src)
    # Euclidean distance to the nearest source, evaluated as the library does
    # ((dx*dx + dy*dy) + dz*dz, then sqrt), so band boundaries agree exactly.
    e = np.full(len(xyz), np.inf)
    for s_ in src:
        dx, dy, dz = (xyz[:, k] - xyz[s_, k] for k in range(3))
        e = np.minimum(e, np.sqrt(dx * dx + dy * dy + dz * dz))
    h = C.mean_edge()
    order, limits = band_order(xyz, src, levels, e, width * h)
    ref, nsw = M.ptp(src, oracle.GPU_COAL | oracle.UNCAPPED, tol=1e-5, toplesets=(None, order, limits))
    # The bare band loop (no check sweeps) is the uncapped loop over the bands.
    plan = P.Plan(C, src, schedule=L.SCHED_EUCLID_BAND, precision=L.PREC_F64, band_width=width, tol=1e-5,
                  band_check=0)
    st = plan.run()
    d = plan.result()
    mx, _ = rel_err(d, ref)
    print(f"[bands f64] {name} w={width} bands {st['n_limits'] - 1} sweeps {st['sweeps']}/{nsw} max {mx:.2e}",
          flush=True)
    assert st["n_limits"] == len(limits)
    assert st["sweeps"] == nsw and st["check_sweeps"] == 0
    assert mx <= 1e-12
    # With the check sweeps (the default) the run ends where no vertex moves
    # by tol.
    plan = P.Plan(C, src, schedule=L.SCHED_EUCLID_BAND, precision=L.PREC_F64, band_width=width, tol=1e-6)
    st = plan.run()
    d = plan.result()
    mxf, _ = rel_err(d, fp)
    print(f"[bands f64] {name} w={width} tol 1e-6 sweeps {st['sweeps']} + {st['check_sweeps']} check sweeps, "
          f"vs fixed point {mxf:.2e}", flush=True)
    assert st["check_sweeps"] >= 1
    assert mxf <= BAND_BAR


@pytest.mark.parametrize("width", [1.0, 4.0])
@pytest.mark.parametrize("sched", [L.SCHED_EUCLID_BAND, L.SCHED_RING_BAND], ids=["euclid", "ringband"])
def test_band_schedule_reaches_fixed_point(case, sched, width):
    name, xyz, f, src, M, C, fp = case
    plan = P.Plan(C, src, schedule=sched, band_width=width, tol=1e-6)
    st = plan.run()
    mx, mean = rel_err(plan.result(), fp)
    print(f"[bands] {name} sched={sched} w={width} sweeps {st['sweeps']} relax {st['relaxations']} max {mx:.3e} "
          f"mean {mean:.3e} kernel {st['kernel_ms']:.2f} ms", flush=True)
    assert mx <= BAND_BAR


def _run(C, src, fp, name, tag, **kw):
    plan = P.Plan(C, src, **kw)
    st = plan.run()
    d = plan.result()
    mx, mean = rel_err(d, fp)
    print(f"[patches] {name} {tag} launches {st['sweeps']} visits {st['patch_visits']} deferred {st['deferred']} "
          f"updates {st['relaxations']} max {mx:.3e} mean {mean:.3e} kernel {st['kernel_ms']:.2f} ms", flush=True)
    plan.close()
    return d, st, mx


@pytest.mark.parametrize("block,li", [(256, 8), (512, 4), (256, 32)], ids=["p256l8", "p512l4", "p256l32"])
def test_patches_reach_fixed_point(case, block, li):
    name, xyz, f, src, M, C, fp = case
    d, st, mx = _run(C, src, fp, name, f"P={block} L={li}", schedule=L.SCHED_PATCHES, block=block, local_iters=li,
                     tol=1e-6)
    assert mx <= 1e-4
    assert st["patch_visits"] > 0 and st["sweeps"] > 0 and st["deferred"] == 0


@pytest.mark.parametrize("pred", ["exact", "zeros", "reversed"])
@pytest.mark.parametrize("sched", [L.SCHED_LEARNED_BAND, L.SCHED_PATCHES], ids=["bands", "patches"])
def test_caller_predictions_only_reorder(case, sched, pred):
    # An exact prediction (the fixed point itself), a useless one and an
    # adversarial one all reach the fixed point.
    name, xyz, f, src, M, C, fp = case
    fin = np.where(np.isfinite(fp), fp, 0.0)
    dhat = {"exact": fin, "zeros": np.zeros_like(fin), "reversed": fin.max() - fin}[pred]
    kw = dict(band_width=1.0) if sched == L.SCHED_LEARNED_BAND else dict(band_width=16.0)
    d, st, mx = _run(C, src, fp, name, f"sched={sched} dhat={pred}", dhat=dhat, schedule=sched, tol=1e-6, **kw)
    assert mx <= (BAND_BAR if sched == L.SCHED_LEARNED_BAND else 1e-4)


def test_launch_modes_agree(case):
    name, xyz, f, src, M, C, fp = case
    dg, _, _ = _run(C, src, fp, name, "graph", schedule=L.SCHED_PATCHES, launch_mode=L.LAUNCH_GRAPH, tol=1e-6)
    ds, _, _ = _run(C, src, fp, name, "stream", schedule=L.SCHED_PATCHES, launch_mode=L.LAUNCH_STREAM, tol=1e-6)
    # In-place updates make the visiting order (hence rounding) run-dependent;
    # both are within tolerance of the same fixed point.
    mx, _ = rel_err(dg, ds)
    assert mx <= 1e-4


def test_patch_options_rejected(case):
    name, xyz, f, src, M, C, fp = case
    for kw in (dict(block=128), dict(local_iters=0), dict(precision=L.PREC_F64), dict(tol=0.0)):
        with pytest.raises(P.GPError) as e:
            P.Plan(C, src, schedule=L.SCHED_PATCHES, **kw)
        assert e.value.status == L.ERR_INVALID
