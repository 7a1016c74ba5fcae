"""Pins of the oracle's update_step (oracle/ptp_impl.h, restating
src/geodesics_ptp.cpp:201-262) by closed forms of the planar-wavefront update.

A triangle (v, a, b) is built as face (v, a, b), so half-edge 0 leaves v with
next -> a and prev -> b, exactly the half-edge update_step receives.
"""
import math

import numpy as np
import pytest


def tri(oracle, v, a, b):
    return oracle.Mesh([v, a, b], [[0, 1, 2]])


def upd(oracle, v, a, b, t0, t1, f32=False):
    if f32:
        gt = np.array([v, a, b], np.float32).reshape(-1)
        return float(oracle.update_step_f32(gt, np.array([0, 1, 2], np.uint32),
                                            np.array([0, t0, t1], np.float32), 0))
    return tri(oracle, v, a, b).update_step(np.array([0.0, t0, t1]), 0)


def rotation(seed):
    q, _ = np.linalg.qr(np.random.default_rng(seed).normal(size=(3, 3)))
    return q


@pytest.mark.parametrize("seed", range(5))
def test_planar_wave_inside_wedge(oracle, seed):
    # Wave t(x) = n.x + c travelling along n; the characteristic through v runs
    # back between a and b, so the update is exact: t(v) = c.
    R = rotation(seed)
    n = np.array([1.0, 0.1, 0.0]) / math.hypot(1.0, 0.1)
    v, a, b = np.zeros(3), np.array([-1.0, 0.5, 0.0]), np.array([-1.0, -0.6, 0.0])
    c = 5.0 + seed
    t0, t1 = float(n @ a + c), float(n @ b + c)
    p = upd(oracle, R @ v, R @ a, R @ b, t0, t1)
    assert abs(p - c) <= 1e-12 * c
    pf = upd(oracle, R @ v, R @ a, R @ b, t0, t1, f32=True)
    assert abs(pf - c) <= 1e-3 * c          # the reference's FP32 build loses digits to cancellation


def test_front_outside_wedge_takes_edge_path(oracle):
    n = np.array([1.0, 0.9, 0.0]) / math.hypot(1.0, 0.9)
    v, a, b = np.zeros(3), np.array([-1.0, 0.5, 0.0]), np.array([-1.0, -0.6, 0.0])
    t0, t1 = float(n @ a + 5), float(n @ b + 5)
    p = upd(oracle, v, a, b, t0, t1)
    assert p == min(t0 + np.linalg.norm(a), t1 + np.linalg.norm(b))
    assert p > 5.0                          # not the planar value


@pytest.mark.parametrize("which", [0, 1])
def test_one_infinite_input(oracle, which):
    v, a, b = np.zeros(3), np.array([-1.0, 0.5, 0.2]), np.array([-1.0, -0.6, 0.1])
    t = [2.0, 3.0]
    t[which] = math.inf
    p = upd(oracle, v, a, b, *t)
    other = b if which == 0 else a
    assert p == t[1 - which] + math.sqrt(float(other @ other))


def test_equilateral_edge_front(oracle):
    # Sources on the whole edge ab (t = 0 at a and b): the front is parallel to
    # ab and v is reached at the triangle's height.
    h = math.sqrt(3) / 2
    p = upd(oracle, [0, 0, 0], [-h, 0.5, 0], [-h, -0.5, 0], 0.0, 0.0)
    assert abs(p - h) <= 1e-15


def test_point_source_at_a_vertex(oracle):
    # t = 0 at a, t = |b - a| at b: the plane wave through a and b runs along
    # ab; for an equilateral triangle its characteristic through v misses the
    # edge, so the update is the edge path |v - a|.
    p = upd(oracle, [0.5, math.sqrt(3) / 2, 0], [0, 0, 0], [1, 0, 0], 0.0, 1.0)
    assert abs(p - 1.0) <= 1e-15


def test_degenerate_triangle_is_nan(oracle):
    p = upd(oracle, [0, 0, 0], [1, 0, 0], [2, 0, 0], 1.0, 2.0)
    assert math.isnan(p)


def test_acute_planar_update_is_causal(oracle):
    # When the planar branch is taken on a triangle acute at v, the front
    # reaches v after both a and b: p > max(t0, t1) (t_k - p = X_k . n with
    # n = c0 X0 + c1 X1, c < 0 and X0 . X1 > 0).
    rng = np.random.default_rng(0)
    seen = 0
    for _ in range(400):
        s = rng.normal(size=3) * 5
        v, a, b = rng.normal(size=(3, 3))
        if (a - v) @ (b - v) <= 0:
            continue
        t0, t1 = float(np.linalg.norm(a - s)), float(np.linalg.norm(b - s))
        p = upd(oracle, v, a, b, t0, t1)
        edge = min(t0 + np.linalg.norm(a - v), t1 + np.linalg.norm(b - v))
        if p != edge:
            seen += 1
            assert p > max(t0, t1)
            assert p <= edge + 1e-12 * edge
    assert seen > 20
