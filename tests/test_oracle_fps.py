"""Oracle farthest-point sampling (farthest_point_sampling_ptp_gpu,
src/cuda/geodesics_ptp.cu:87-172), pinned by what the definition fixes."""
import numpy as np
import pytest

from gproshan_amd import meshgen as G


def flat_grid(n):
    s = np.linspace(0.0, 1.0, n)
    y, x = np.meshgrid(s, s, indexing="ij")
    xyz = np.stack([x.ravel(), y.ravel(), np.zeros(n * n)], 1)
    flip = np.zeros((n - 1, n - 1), bool)
    return xyz, G.grid_faces(n, n, flip)


def test_first_sample_is_the_opposite_corner(oracle):
    # From corner (0,0) of a flat unit square the farthest vertex is (1,1):
    # every other vertex is at least ~0.29 closer in the plane.
    n = 17
    xyz, f = flat_grid(n)
    M = oracle.Mesh(xyz, f)
    s, _ = M.fps([0], 2)
    assert s.tolist() == [0, n * n - 1]


def test_each_sample_is_the_farthest_of_the_previous_field(oracle):
    # Round k appends the first index of the largest distance of the field of
    # samples[:k] (the GPU plain driver, computed independently here).
    xyz, f = G.height_field(24, 3)
    M = oracle.Mesh(xyz, f)
    s, _ = M.fps([5], 8)
    for k in range(1, len(s)):
        d, _ = M.ptp(s[:k], oracle.GPU_PLAIN)
        assert s[k] == int(np.argmax(np.abs(d)))
        assert s[k] not in s[:k]


def test_budget_and_radius(oracle):
    xyz, f = G.icosphere(2, 0.01, 4)             # 162 vertices
    M = oracle.Mesh(xyz, f)
    s, r = M.fps([0], 10 ** 9)                   # n >= V: V/2 samples in all
    assert len(s) == len(xyz) // 2 and len(set(s.tolist())) == len(s)
    assert 0 < r < np.inf                        # the last round's maximum
    s1, r1 = M.fps([0], 50, radio=10.0)          # radius above every distance: one round
    assert len(s1) == 2 and r1 < 10.0
    s2, r2 = M.fps([0, 1], 2)                    # nothing to add
    assert s2.tolist() == [0, 1] and r2 == np.inf
