"""Oracle cluster labels (the reference GPU drivers' rule,
src/cuda/geodesics_ptp.cu:257-282), pinned by what the definition fixes."""
import numpy as np

from gproshan_amd import meshgen as G


def test_one_source_labels_everything_one(oracle):
    xyz, f = G.height_field(32, 1)
    M = oracle.Mesh(xyz, f)
    d, cl, _ = M.ptp_clusters([100])
    assert np.all(cl == 1)
    ref, _ = M.ptp([100], oracle.GPU_COAL)
    assert np.array_equal(d, ref)                  # labels do not change the distances


def test_labels_follow_the_nearer_source(oracle):
    # Away from the bisector (margin > 10 % of the field) the label is the
    # source whose own single-source field is smaller.
    xyz, f = G.height_field(48, 2)
    M = oracle.Mesh(xyz, f)
    src = np.array([50, 48 * 48 - 60], np.uint32)
    _, cl, _ = M.ptp_clusters(src)
    d1, _ = M.ptp(src[:1], oracle.GPU_COAL)
    d2, _ = M.ptp(src[1:], oracle.GPU_COAL)
    clear = np.abs(d1 - d2) > 0.1 * max(d1.max(), d2.max())
    want = np.where(d1 < d2, 1, 2)
    assert clear.sum() > 1000
    assert np.array_equal(cl[clear], want[clear])
    assert set(np.unique(cl).tolist()) == {1, 2}


def test_plain_and_coalesced_agree_and_unreached_is_nil(oracle):
    a, fa = G.height_field(20, 5)
    b, fb = G.icosphere(1, 0.0, 1)
    xyz = np.concatenate([a, b * 0.1 + 5.0])
    f = np.concatenate([fa.reshape(-1, 3), fb.reshape(-1, 3) + len(a)])
    M = oracle.Mesh(xyz, f)
    src = np.array([7, 333], np.uint32)
    dp, cp, _ = M.ptp_clusters(src, oracle.GPU_PLAIN)
    dc, cc, _ = M.ptp_clusters(src, oracle.GPU_COAL)
    assert np.array_equal(cp, cc) and np.array_equal(dp, dc)
    assert np.all(cp[len(a):] == 0xFFFFFFFF)        # the island no source reaches
    assert cp[7] == 1 and cp[333] == 2
