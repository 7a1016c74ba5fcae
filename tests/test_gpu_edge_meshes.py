"""GPU parity on meshes the benchmark workloads do not contain.

  * two components (a height field and an icosphere): the component without a
    source stays INF, and with a source in each both fronts run in one window;
  * a height field with holes cut out (internal boundaries, open fans inside);
  * 300 sources (ring 0 spans several 64-vertex chunks);
  * collapsed vertices (zero-length edges): the reference's update divides by
    its determinant q00 q11 - q01 q10, exactly 0 there, and returns NaN, which
    loses every comparison (src/geodesics_ptp.cpp:201-262); the FP32 path's
    tables mark such corners (det = NaN, FP64 fallback) so that they lose too
    -- before they did, the FP32 field of this mesh was 6.2e-2 off;
  * tiny meshes: one triangle, a quad, and a mesh whose every vertex is a
    source (no sweep runs);
  * coordinates far from the origin (3e5 offset) and at a micro scale (1e-6):
    the FP32 path works on FP64 differences in units of the mean edge;
  * strips of thin quads (every eighth column 20x narrower): right-angle
    corners whose two edges differ 20x in length, which only the edge-ratio
    test sends to the FP64 update;
  * non-manifold input (a bowtie vertex, an edge with three triangles), where
    the half-edge tables follow the reference's (NIL twins and EVT) and so does
    what propagates.

Each mesh: GPU toplesets bit-identical to the oracle's; the reference schedule
in FP64 bit-identical to the oracle's restatement (coalesced and sorted[]
drivers, same sweep count), with host- and device-built rings alike; the
uncapped loop in FP64 likewise; FP32 within 1e-4 of the FP64 oracle; the
dataflow launch bit-identical to the graph path.
"""
import numpy as np
import pytest

from conftest import rel_err
from gproshan_amd import _lib as L
from gproshan_amd import geodesics as P
from gproshan_amd import meshgen as G

pytestmark = pytest.mark.gpu


def two_components():
    a, fa = G.height_field(64, 31)
    b, fb = G.icosphere(3, 0.01, 2)
    b = b + np.array([5.0, 0.0, 0.0])
    return np.vstack([a, b]), np.vstack([fa, fb + len(a)]).astype(np.uint32)


def holes():
    n = 96
    xyz, f = G.height_field(n, 32)
    # Drop the faces whose first vertex lies in one of three square windows.
    r, c = f[:, 0] // n, f[:, 0] % n
    cut = np.zeros(len(f), bool)
    for r0, c0, s in ((20, 20, 8), (50, 60, 12), (70, 15, 6)):
        cut |= (r >= r0) & (r < r0 + s) & (c >= c0) & (c < c0 + s)
    keep = f[~cut]
    used = np.unique(keep)
    remap = np.full(len(xyz), -1, np.int64)
    remap[used] = np.arange(len(used))
    return xyz[used], remap[keep].astype(np.uint32)


def collapsed():
    n = 64
    xyz, f = G.height_field(n, 33)
    xyz = xyz.copy()
    for v in (n * 10 + 10, n * 30 + 41, n * 50 + 7):
        xyz[v] = xyz[v + 1]          # the vertex moves onto its right neighbour
    return xyz, f


def strips():
    n = 96
    xyz, f = G.height_field(n, 38)
    xyz = xyz.copy()
    gap = np.where(np.arange(n - 1) % 8 == 7, 0.05, 1.0)
    xs = np.concatenate([[0.0], np.cumsum(gap)])
    xyz[:, 0] = (xs / xs[-1])[np.arange(len(xyz)) % n]
    xyz[:, 2] *= 0.02          # nearly flat, so a thin quad's short edge stays short
    return xyz, f


def one_triangle():
    return np.array([[0, 0, 0], [1, 0, 0], [0.2, 0.9, 0.1]], float), np.array([[0, 1, 2]], np.uint32)


def bowtie():
    """Two triangles sharing only vertex 0 (non-manifold: the reference's
    che leaves EVT[0] = NIL and propagates nothing through vertex 0)."""
    return (np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [-1, 0, 0], [0, -1, 0]], float),
            np.array([[0, 1, 2], [0, 3, 4]], np.uint32))


def book3():
    """Three triangles on one edge (a non-manifold edge)."""
    return (np.array([[0, 0, 0], [1, 0, 0], [0.5, 1, 0], [0.5, -1, 0], [0.5, 0, 1]], float),
            np.array([[0, 1, 2], [1, 0, 3], [0, 1, 4]], np.uint32))


def quad():
    return (np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0.2], [0, 1, 0]], float),
            np.array([[0, 1, 2], [0, 2, 3]], np.uint32))


CASES = [
    ("two_components_one_source", two_components, lambda nv: [100]),
    ("two_components_two_sources", two_components, lambda nv: [100, nv - 5]),
    ("holes", holes, lambda nv: [nv // 2]),
    ("sources300", lambda: G.height_field(128, 34), lambda nv: G.draw_sources(np.random.default_rng(34), nv, 300)),
    ("collapsed", collapsed, lambda nv: [n for n in (64 * 32 + 32,)]),
    ("strips", strips, lambda nv: [96 * 48 + 20]),
    ("one_triangle", one_triangle, lambda nv: [0]),
    ("quad", quad, lambda nv: [1]),
    ("all_sources", lambda: G.height_field(8, 35), lambda nv: list(range(nv))),
    ("far_offset", lambda: (G.height_field(96, 36)[0] + np.array([3.0e5, -2.0e5, 1.0e5]), G.height_field(96, 36)[1]),
     lambda nv: [4000]),
    ("micro_scale", lambda: (G.height_field(96, 37)[0] * 1e-6, G.height_field(96, 37)[1]), lambda nv: [5000]),
    ("bowtie", bowtie, lambda nv: [1]),
    ("book3", book3, lambda nv: [2]),
]


@pytest.fixture(scope="module", params=CASES, ids=[c[0] for c in CASES])
def case(request, oracle):
    name, gen, pick = request.param
    xyz, f = gen()
    src = np.array(sorted(set(int(s) for s in pick(len(xyz)))), np.uint32)
    print(f"\n[case] {name} V={len(xyz)} F={len(f)} sources={len(src)}", flush=True)
    return name, xyz, f, src, oracle.Mesh(xyz, f), P.che(xyz, f)


def _same(a, b):
    return np.array_equal(np.asarray(a, np.float64), np.asarray(b, np.float64), equal_nan=True)


def test_toplesets(case):
    name, xyz, f, src, M, C = case
    ref = M.toplesets(src)
    got = C.compute_toplesets(src, gpu=True)
    for a, b, what in zip(got, ref, ("toplesets", "sorted", "limits")):
        assert np.array_equal(a, b), what


@pytest.mark.parametrize("coal", [True, False], ids=["coalesce", "plain"])
def test_f64_reference_schedule(case, oracle, coal):
    name, xyz, f, src, M, C = case
    ref, nsw = M.ptp(src, oracle.GPU_COAL if coal else oracle.GPU_PLAIN)
    d, _, st = P.ptp(C, src, precision=L.PREC_F64, coalesce=coal)
    print(f"  f64 coal={coal}: sweeps {st['sweeps']}/{nsw}", flush=True)
    assert st["sweeps"] == nsw
    assert _same(d, ref)


def test_f64_device_built(case, oracle):
    name, xyz, f, src, M, C = case
    ref, nsw = M.ptp(src, oracle.GPU_COAL)
    plan = P.Plan(C, src, precision=L.PREC_F64, device_setup=True)
    st = plan.run()
    assert st["sweeps"] == nsw
    assert _same(plan.result(), ref)
    plan.close()


def test_f64_uncapped(case, oracle):
    name, xyz, f, src, M, C = case
    ref, nsw = M.ptp(src, oracle.GPU_COAL | oracle.UNCAPPED, tol=1e-5)
    d, _, st = P.ptp(C, src, precision=L.PREC_F64, tol=1e-5, uncapped=True)
    assert st["sweeps"] == nsw
    assert _same(d, ref)


def test_f32_within_bar(case, oracle):
    name, xyz, f, src, M, C = case
    ref, _ = M.ptp(src, oracle.GPU_COAL)
    d, _, st = P.ptp(C, src, precision=L.PREC_F32, fp32_geometry=L.GEO_FP32)
    mx, mean = rel_err(d, ref)
    print(f"  f32: max {mx:.2e} mean {mean:.2e} f64 updates {st['f64_updates']}", flush=True)
    assert mx <= 1e-4


def test_flow_matches_graph(case):
    name, xyz, f, src, M, C = case
    out = {}
    for mode in (L.LAUNCH_GRAPH, L.LAUNCH_FLOW):
        plan = P.Plan(C, src, precision=L.PREC_F32, fp32_geometry=L.GEO_FP32, uncapped=True, tol=1e-6,
                      launch_mode=mode)
        st = plan.run()
        out[mode] = (st["sweeps"], plan.result())
        plan.close()
    assert out[L.LAUNCH_GRAPH][0] == out[L.LAUNCH_FLOW][0]
    assert _same(out[L.LAUNCH_GRAPH][1], out[L.LAUNCH_FLOW][1])


def test_batch_on_edge_meshes(oracle):
    """The batched solver on the holed, collapsed and two-component meshes in
    one launch: every field within 1e-4 of the FP64 oracle (sorted[] driver,
    variant GPU_PLAIN) and equal to the single-field FP32 solve."""
    gens = [("holes", holes), ("collapsed", collapsed), ("two_components", two_components)]
    meshes, orc, jobs = [], [], []
    for k, (name, gen) in enumerate(gens):
        xyz, f = gen()
        meshes.append(P.che(xyz, f))
        orc.append(oracle.Mesh(xyz, f))
        for s in (0, len(xyz) // 3, len(xyz) - 1):
            jobs.append((k, s))
    B = P.Batch(meshes, jobs, block=1024)
    B.run()
    for q, (k, s) in enumerate(jobs):
        d, sw = B.result(q)
        ref, nsw = orc[k].ptp(np.array([s], np.uint32), oracle.GPU_PLAIN)
        mx, _ = rel_err(d, ref)
        single, _, st = P.ptp(meshes[k], [s], precision=L.PREC_F32, fp32_geometry=L.GEO_FP32, coalesce=False)
        print(f"  batch {gens[k][0]} source {s}: sweeps {sw}/{nsw} max {mx:.2e}", flush=True)
        assert mx <= 1e-4
        assert sw == st["sweeps"]
        assert _same(d, single)
    B.close()


@pytest.mark.parametrize("coal", [True, False], ids=["coalesce", "plain"])
def test_cluster_labels_f64(case, oracle, coal):
    """Cluster labels (src/cuda/geodesics_ptp.cu:257-282) in FP64: labels and
    distances identical to the oracle's, NIL where unreached."""
    name, xyz, f, src, M, C = case
    dref, cref, nsw = M.ptp_clusters(src, oracle.GPU_COAL if coal else oracle.GPU_PLAIN)
    plan = P.Plan(C, src, precision=L.PREC_F64, coalesce=coal, clusters=True)
    st = plan.run()
    assert st["sweeps"] == nsw
    assert _same(plan.result(), dref)
    assert np.array_equal(plan.clusters(), cref)
    plan.close()
