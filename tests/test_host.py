"""CPU tests of libgproshan_ptp: the C ABI loads and exports every declared
symbol, the host CHE builder / toplesets / FP64 host solver match the oracle
bit for bit, and error paths return status codes (no compute on a GPU)."""
import ctypes
import os
import re

import numpy as np
import pytest

from gproshan_amd import _lib as L
from gproshan_amd import geodesics as P
from gproshan_amd import meshgen as G

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "gproshan_ptp.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gp_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(L.LIB_PATH)
    decl = declared_functions()
    assert len(decl) >= 20
    missing = [name for name in decl if not hasattr(lib, name)]
    assert not missing, missing
    assert sorted(L.EXPORTS) == decl


MESHES = [
    ("ico4", lambda: G.icosphere(4, 0.01, 0), [0]),
    ("hf64", lambda: G.height_field(64, 3), [100, 2000, 4000]),
    ("aniso24", lambda: G.anisotropic_grid(24, 100.0, 0.4, 1), [300]),
    ("bowtie", lambda: (np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [-1, 0, 0], [-1, -1, 0]], float),
                        np.array([[0, 1, 2], [0, 3, 4]], np.uint32)), [1]),
]


@pytest.fixture(scope="module", params=MESHES, ids=[m[0] for m in MESHES])
def pair(request, oracle):
    name, gen, src = request.param
    xyz, f = gen()
    return oracle.Mesh(xyz, f), P.che(xyz, f), np.array(src, np.uint32)


def test_che_tables_match_oracle(pair):
    M, C, _ = pair
    vt, ot, evt = C.tables()
    assert np.array_equal(vt, M.vt) and np.array_equal(ot, M.ot) and np.array_equal(evt, M.evt)
    assert C.n_edges() == M.ne and C.is_manifold() == M.manifold
    for v in range(0, M.nv, max(1, M.nv // 97)):
        assert np.array_equal(C.link(v), M.link(v))


def test_host_toplesets_match_oracle(pair):
    M, C, src = pair
    for a, b in zip(C.compute_toplesets(src), M.toplesets(src)):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("k", [0, 1, 2, 5])
def test_host_toplesets_level_cut_match_oracle(pair, k):
    M, C, src = pair
    for a, b in zip(C.compute_toplesets(src, k=k), M.toplesets(src, k=k)):
        assert np.array_equal(a, b)


def test_toplesets_level_cut_by_hand():
    # A strip of four unit squares, vertices 0-4 at y = 0 and 5-9 at y = 1,
    # each square split along its (i, i + 6) diagonal.  From source 0:
    # level 1 = {1, 6, 5}, level 2 = {2, 7}, level 3 = {3, 8}, ...  With the
    # cut k = 1 the walk stops at the first level-2 vertex, after the level-1
    # vertices have discovered {2, 7}: those stay in sorted, at level 2, and in
    # the last ring (src/che.cpp:572-575, 592).
    xyz = np.array([[x, y, 0.0] for y in (0, 1) for x in range(5)])
    f = []
    for i in range(4):
        f += [[i, i + 1, i + 6], [i, i + 6, i + 5]]
    C = P.che(xyz, np.array(f, np.uint32))
    lev, srt, lim = C.compute_toplesets([0], k=1)
    assert lim.tolist() == [0, 1, 6]
    assert sorted(srt[1:4].tolist()) == [1, 5, 6] and sorted(srt[4:6].tolist()) == [2, 7]
    assert (srt[6:] == L.NIL).all()
    assert lev[[0, 1, 5, 6, 2, 7]].tolist() == [0, 1, 1, 1, 2, 2] and (lev[[3, 4, 8, 9]] == L.NIL).all()
    lev0, srt0, lim0 = C.compute_toplesets([0])
    assert lim0.tolist() == [0, 1, 4, 6, 8, 10]


def test_host_ptp_cpu_bit_exact(pair, oracle):
    M, C, src = pair
    ts = C.compute_toplesets(src)
    assert np.array_equal(P.parallel_toplesets_propagation_cpu(C, src, ts), M.ptp(src, oracle.CPU_PLAIN)[0])
    assert np.array_equal(P.parallel_toplesets_propagation_coalescence_cpu(C, src, ts),
                          M.ptp(src, oracle.CPU_COAL)[0])


def test_frontend_ptp_cpu(pair, oracle):
    M, C, src = pair
    g = P.geodesics(C, src, P.geodesics.PTP_CPU)
    assert np.array_equal(g.dist, M.ptp(src, oracle.CPU_COAL)[0])
    srt = M.toplesets(src)[1]
    assert [g(i) for i in range(len(srt))] == srt.tolist()
    assert g.n_sorted_index() == 0
    with pytest.raises(P.GPError):
        g.farthest()
    d = g.dist.copy()
    g.normalize()
    assert np.array_equal(g.dist, oracle.normalize_ptp(d))


def test_error_paths():
    xyz, f = G.icosphere(1, 0.0, 0)
    C = P.che(xyz, f)
    with pytest.raises(P.GPError) as e:
        C.compute_toplesets([3, 3])
    assert e.value.status == L.ERR_DUPLICATE_SOURCE
    with pytest.raises(P.GPError) as e:
        C.compute_toplesets([len(xyz)])
    assert e.value.status == L.ERR_INVALID
    with pytest.raises(P.GPError) as e:
        P.che(xyz, np.array([[0, 1, len(xyz)]], np.uint32))
    assert e.value.status == L.ERR_INVALID
    with pytest.raises(P.GPError):
        P.geodesics(C, [], P.geodesics.PTP_CPU)
    with pytest.raises(P.GPError):
        P.geodesics(C, [0], P.geodesics.FM)
    # empty source list: toplesets returns no rings, as the reference does
    top, srt, lim = C.compute_toplesets([])
    assert len(lim) == 0
    # batch jobs are validated before any device work
    for jobs, status in (([(1, 0)], L.ERR_INVALID), ([(0, len(xyz))], L.ERR_INVALID)):
        with pytest.raises(P.GPError) as e:
            P.Batch([C], jobs)
        assert e.value.status == status
    with pytest.raises(P.GPError) as e:
        P.Batch([C], [(0, 0)], tol=0.0)
    assert e.value.status == L.ERR_INVALID


def test_device_probe_reports_missing_gpu():
    # In a container without a GPU the probe fails loudly with a status code.
    buf = ctypes.create_string_buffer(64)
    rc = L.lib().gp_device_check(buf, 64)
    if rc != L.OK:
        assert rc in (L.ERR_NO_DEVICE, L.ERR_HIP)
        assert L.lib().gp_last_error()
