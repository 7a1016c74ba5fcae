"""Pins of the oracle's compact half-edge construction and toplesets
(oracle/che.c) by hand-derived tables and graph invariants.

The reference ships no tests or fixtures for this path; the tables below were
derived by hand from the construction rules of src/che.cpp:1295-1362 (OT:
first unpaired reversed half-edge in ascending scan; EVT: largest outgoing
half-edge, replaced by the outgoing border half-edge, NIL for two border fans),
for_star (include/che.h:10), che::link (src/che.cpp:102-112) and
che::compute_toplesets (src/che.cpp:546-593).
"""
import numpy as np
import pytest
import scipy.sparse as sp
from scipy.sparse.csgraph import shortest_path

from gproshan_amd import meshgen as G

NIL = 0xFFFFFFFF


def nxt(he):
    return 3 * (he // 3) + (he + 1) % 3


def prv(he):
    return 3 * (he // 3) + (he + 2) % 3


TETRA_XYZ = [[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]]
TETRA_F = [[0, 2, 1], [0, 1, 3], [0, 3, 2], [1, 2, 3]]

GRID2_XYZ = [[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0]]
GRID2_F = [[0, 1, 3], [0, 3, 2]]

BOWTIE_XYZ = [[0, 0, 0], [1, 0, 0], [1, 1, 0], [-1, 0, 0], [-1, -1, 0]]
BOWTIE_F = [[0, 1, 2], [0, 3, 4]]


def test_tetrahedron_tables(oracle):
    m = oracle.Mesh(TETRA_XYZ, TETRA_F)
    assert m.vt.tolist() == [0, 2, 1, 0, 1, 3, 0, 3, 2, 1, 2, 3]
    assert m.ot.tolist() == [8, 9, 3, 2, 11, 6, 5, 10, 0, 1, 7, 4]
    assert m.evt.tolist() == [6, 9, 10, 11]
    assert m.ne == 6 and m.manifold
    assert m.star(0).tolist() == [6, 0, 3]
    assert [int(m.vt[h]) for h in m.link(0)] == [3, 2, 1]
    lev, srt, lim = m.toplesets([0])
    assert srt.tolist() == [0, 3, 2, 1]
    assert lev.tolist() == [0, 1, 1, 1]
    assert lim.tolist() == [0, 1, 4]


def test_open_grid_tables(oracle):
    m = oracle.Mesh(GRID2_XYZ, GRID2_F)
    assert m.ot.tolist() == [NIL, NIL, 3, 2, NIL, NIL]
    assert m.evt.tolist() == [0, 1, 5, 4]              # border rule moved EVT[0] from 3 to 0
    assert m.manifold
    assert [int(m.vt[h]) for h in m.link(0)] == [1, 3, 2]
    assert [int(m.vt[h]) for h in m.link(3)] == [2, 0, 1]
    assert m.toplesets([0])[1].tolist() == [0, 1, 3, 2]
    assert m.toplesets([3])[1].tolist() == [3, 2, 0, 1]


def test_bowtie_nonmanifold(oracle):
    m = oracle.Mesh(BOWTIE_XYZ, BOWTIE_F)
    assert m.ot.tolist() == [NIL] * 6
    assert m.evt.tolist() == [NIL, 1, 2, 4, 5]
    assert not m.manifold
    assert m.star(0).size == 0
    lev, srt, lim = m.toplesets([0])
    assert lim.tolist() == [0, 1]
    lev, srt, lim = m.toplesets([1])
    assert srt[:3].tolist() == [1, 2, 0] and lim.tolist() == [0, 1, 3]
    assert lev[3] == NIL and lev[4] == NIL and srt[3] == NIL


@pytest.mark.parametrize("gen", [lambda: G.icosphere(2, 0.01, 3), lambda: G.height_field(17, 5),
                                 lambda: G.anisotropic_grid(12, 100.0, 0.4, 2)])
def test_che_invariants(oracle, gen):
    xyz, f = gen()
    m = oracle.Mesh(xyz, f)
    he = np.arange(m.nhe)
    paired = m.ot != NIL
    o = m.ot[paired]
    assert np.array_equal(m.ot[o], he[paired])                       # OT is an involution
    assert np.array_equal(m.vt[o], m.vt[nxt(he[paired])])            # opposite runs backwards
    assert np.array_equal(m.vt[nxt(o)], m.vt[he[paired]])
    # EVT: largest outgoing half-edge for interior vertices, the outgoing
    # border half-edge for border vertices.
    for v in range(m.nv):
        out = he[m.vt == v]
        border = out[~paired[out]]
        if len(border) == 0:
            assert m.evt[v] == out.max()
        elif len(border) == 1:
            assert m.evt[v] == border[0]
    # Euler characteristic: closed sphere 2, disc 1.
    chi = m.nv - m.ne + len(f)
    assert chi == (2 if paired.all() else 1)


def _edge_graph(xyz, f):
    f = np.asarray(f, dtype=np.int64)
    e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
    n = len(xyz)
    return sp.coo_matrix((np.ones(len(e)), (e[:, 0], e[:, 1])), shape=(n, n)).tocsr()


@pytest.mark.parametrize("gen,src", [(lambda: G.icosphere(3, 0.01, 1), [0]),
                                     (lambda: G.height_field(33, 2), [100, 700]),
                                     (lambda: G.height_field(20, 9), [0, 399, 210])])
def test_toplesets_properties(oracle, gen, src):
    xyz, f = gen()
    m = oracle.Mesh(xyz, f)
    lev, srt, lim = m.toplesets(src)
    hops = shortest_path(_edge_graph(xyz, f), unweighted=True, directed=False, indices=src).min(axis=0)
    assert np.array_equal(lev.astype(float), hops)                      # levels are hop distances
    n = lim[-1]
    assert sorted(srt[:n].tolist()) == sorted(np.flatnonzero(np.isfinite(hops)).tolist())
    assert srt[:len(src)].tolist() == list(src)                        # sources first, in order
    counts = np.bincount(lev[srt[:n]])
    assert np.array_equal(lim, np.concatenate([[0], np.cumsum(counts)]))
    # Discovery order: within a ring, the first discoverer's position (the
    # smallest position of a neighbour in the previous ring) never decreases.
    pos = np.full(m.nv, -1)
    pos[srt[:n]] = np.arange(n)
    adj = _edge_graph(xyz, f)
    adj = (adj + adj.T).tolil()
    for r in range(1, len(lim) - 1):
        ring = srt[lim[r]:lim[r + 1]]
        firsts = [min(pos[u] for u in adj.rows[v] if lev[u] == r - 1) for v in ring]
        assert all(a <= b for a, b in zip(firsts, firsts[1:]))


def test_toplesets_rejects_duplicate_sources(oracle):
    m = oracle.Mesh(TETRA_XYZ, TETRA_F)
    with pytest.raises(ValueError):
        m.toplesets([1, 1])


def test_generator_sizes():
    xyz, f = G.icosphere(5, 0.01, 0)
    assert xyz.shape == (10242, 3) and f.shape == (20480, 3)
    xyz, f = G.height_field(64, 0)
    assert xyz.shape == (4096, 3) and f.shape == (2 * 63 * 63, 3)
    xyz, f = G.anisotropic_grid(32, 100.0, 0.4, 0)
    assert xyz.shape == (1024, 3) and np.all(xyz[:, 2] == 0)
    # determinism
    a, _ = G.height_field(32, 7)
    b, _ = G.height_field(32, 7)
    assert np.array_equal(a, b)
