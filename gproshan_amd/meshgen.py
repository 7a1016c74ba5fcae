"""Seeded synthetic meshes for the PTP workloads.

Every generator is deterministic in its seed (numpy PCG64 via
``np.random.default_rng``) and returns ``(xyz float64[V,3], faces uint32[F,3])``
with consistently counter-clockwise faces, which the compact half-edge builder
needs to pair opposite half-edges (the reference pairs a half-edge only with a
reversed one, src/che.cpp:1314-1331).

Workloads:
  * ``icosphere(level, sigma, seed)``: icosahedron, ``level`` midpoint
    subdivisions projected to the unit sphere, radius ``1 + N(0, sigma)`` per
    vertex (level 5: 10,242 vertices, 20,480 faces).
  * ``height_field(n, seed)``: ``n x n`` grid on the unit square, vertex id
    ``row * n + col``, each quad split along a Bernoulli(0.5) diagonal,
    ``z = sum_k a_k sin(2 pi f_k (x cos th_k + y sin th_k) + ph_k)`` with
    ``a ~ U(0, 0.05)``, ``f ~ U(1, 16)``, ``th, ph ~ U(0, 2 pi)``, 8 terms
    (n = 1024: 1,048,576 vertices, 2,093,058 faces).
  * ``anisotropic_grid(n, stretch, jitter, seed)``: planar ``n x n`` grid with
    cells ``stretch`` times wider in x than in y and every vertex moved by
    ``U(-jitter, jitter)`` cells in x and y, giving obtuse and skinny
    triangles (n = 2048, stretch = 100, jitter = 0.4: 4,194,304 vertices).
  * ``draw_sources(rng, nv, k)``: ``k`` distinct vertices.

The draw order inside each generator is fixed and documented in the code so
that any reimplementation can reproduce the meshes bit for bit.
"""
from __future__ import annotations

import numpy as np

__all__ = ["icosphere", "height_field", "anisotropic_grid", "draw_sources", "grid_faces"]


def _icosahedron():
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v = np.array([
        [-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0],
        [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
        [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1],
    ], dtype=np.float64)
    f = np.array([
        [0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11],
        [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8],
        [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
        [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1],
    ], dtype=np.int64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    return v, f


def icosphere(level: int = 5, sigma: float = 0.01, seed: int = 0, radius: float = 1.0):
    """Displaced icosphere.  Subdivision splits face (a, b, c) into
    (a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca); new midpoint vertices
    are numbered in order of first creation.  After projection to the unit
    sphere each vertex is scaled by ``radius * (1 + eps)``, eps ~ N(0, sigma),
    drawn once for all vertices in vertex order."""
    v, f = _icosahedron()
    verts = [row for row in v]
    for _ in range(level):
        cache: dict[tuple[int, int], int] = {}

        def mid(a: int, b: int) -> int:
            key = (a, b) if a < b else (b, a)
            idx = cache.get(key)
            if idx is None:
                p = verts[a] + verts[b]
                p = p / np.linalg.norm(p)
                idx = len(verts)
                verts.append(p)
                cache[key] = idx
            return idx

        nf = np.empty((4 * len(f), 3), dtype=np.int64)
        for k, (a, b, c) in enumerate(f):
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf[4 * k + 0] = (a, ab, ca)
            nf[4 * k + 1] = (b, bc, ab)
            nf[4 * k + 2] = (c, ca, bc)
            nf[4 * k + 3] = (ab, bc, ca)
        f = nf
    xyz = np.asarray(verts, dtype=np.float64)
    xyz /= np.linalg.norm(xyz, axis=1, keepdims=True)
    rng = np.random.default_rng(seed)
    eps = rng.normal(0.0, sigma, size=len(xyz)) if sigma > 0 else np.zeros(len(xyz))
    xyz *= (radius * (1.0 + eps))[:, None]
    return xyz, f.astype(np.uint32)


def grid_faces(n_rows: int, n_cols: int, flip: np.ndarray) -> np.ndarray:
    """Faces of a grid with vertex id ``row * n_cols + col``.  Quad (r, c) has
    corners a=(r,c), b=(r,c+1), c=(r+1,c+1), d=(r+1,c); ``flip[r, c]`` False
    splits it as (a,b,c),(a,c,d), True as (a,b,d),(b,c,d).  Both triangles are
    counter-clockwise seen from +z.  Quads are emitted row-major."""
    r, c = np.meshgrid(np.arange(n_rows - 1), np.arange(n_cols - 1), indexing="ij")
    a = (r * n_cols + c).ravel()
    b = a + 1
    cc = a + n_cols + 1
    d = a + n_cols
    fl = flip.ravel()
    t0 = np.where(fl[:, None], np.stack([a, b, d], 1), np.stack([a, b, cc], 1))
    t1 = np.where(fl[:, None], np.stack([b, cc, d], 1), np.stack([a, cc, d], 1))
    faces = np.empty((2 * len(a), 3), dtype=np.uint32)
    faces[0::2] = t0
    faces[1::2] = t1
    return faces


def height_field(n: int = 1024, seed: int = 0, n_terms: int = 8):
    """Height field on [0,1]^2.  Draw order: a[8], f[8], theta[8], phi[8],
    then the (n-1)^2 diagonal flips (row-major)."""
    rng = np.random.default_rng(seed)
    amp = rng.uniform(0.0, 0.05, n_terms)
    freq = rng.uniform(1.0, 16.0, n_terms)
    theta = rng.uniform(0.0, 2.0 * np.pi, n_terms)
    phase = rng.uniform(0.0, 2.0 * np.pi, n_terms)
    flip = rng.random((n - 1, n - 1)) < 0.5
    s = np.linspace(0.0, 1.0, n)
    y, x = np.meshgrid(s, s, indexing="ij")
    z = np.zeros_like(x)
    for k in range(n_terms):
        z += amp[k] * np.sin(2.0 * np.pi * freq[k] * (x * np.cos(theta[k]) + y * np.sin(theta[k])) + phase[k])
    xyz = np.stack([x.ravel(), y.ravel(), z.ravel()], 1)
    return xyz, grid_faces(n, n, flip)


def height_field_rect(n_rows: int, n_cols: int, seed: int = 0, n_terms: int = 8):
    """Rectangular variant of ``height_field`` (same draws; x spans
    ``[0, (n_cols-1)/(n_rows-1)]`` so cells stay square)."""
    rng = np.random.default_rng(seed)
    amp = rng.uniform(0.0, 0.05, n_terms)
    freq = rng.uniform(1.0, 16.0, n_terms)
    theta = rng.uniform(0.0, 2.0 * np.pi, n_terms)
    phase = rng.uniform(0.0, 2.0 * np.pi, n_terms)
    flip = rng.random((n_rows - 1, n_cols - 1)) < 0.5
    h = 1.0 / max(n_rows - 1, 1)
    y, x = np.meshgrid(np.arange(n_rows) * h, np.arange(n_cols) * h, indexing="ij")
    z = np.zeros_like(x)
    for k in range(n_terms):
        z += amp[k] * np.sin(2.0 * np.pi * freq[k] * (x * np.cos(theta[k]) + y * np.sin(theta[k])) + phase[k])
    xyz = np.stack([x.ravel(), y.ravel(), z.ravel()], 1)
    return xyz, grid_faces(n_rows, n_cols, flip)


def anisotropic_grid(n: int = 2048, stretch: float = 100.0, jitter: float = 0.4, seed: int = 0):
    """Planar jittered grid, cell size (stretch*h, h) with h = 1/(n-1).  Draw
    order: x jitter [n*n], y jitter [n*n], then the (n-1)^2 diagonal flips."""
    rng = np.random.default_rng(seed)
    h = 1.0 / (n - 1)
    jx = rng.uniform(-jitter, jitter, n * n)
    jy = rng.uniform(-jitter, jitter, n * n)
    flip = rng.random((n - 1, n - 1)) < 0.5
    r, c = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    x = (c.ravel() + jx) * h * stretch
    y = (r.ravel() + jy) * h
    xyz = np.stack([x, y, np.zeros_like(x)], 1)
    return xyz, grid_faces(n, n, flip)


def polar_disk(n_rings: int = 40, n_sect: int = 24, seed: int = 0):
    """Disk around a centre vertex of valence ``n_sect``: ``n_rings`` rings of
    ``n_sect`` vertices (odd rings rotated half a sector, radii jittered by up
    to 10 % of the ring spacing, heights by +-0.025), quads split along
    alternating diagonals.  The centre's fan is longer than the register-held
    slots of the sweep kernels when ``n_sect`` > 10.  Draw order: per ring,
    per sector, radius jitter then height."""
    rng = np.random.default_rng(seed)
    pts = [[0.0, 0.0, 0.0]]
    for r in range(1, n_rings + 1):
        for k in range(n_sect):
            a = 2 * np.pi * (k + 0.5 * (r % 2)) / n_sect
            rad = r + 0.1 * rng.uniform(-1, 1)
            pts.append([rad * np.cos(a), rad * np.sin(a), 0.05 * (rng.uniform() - 0.5)])

    def vid(r, k):
        return 1 + (r - 1) * n_sect + (k % n_sect)
    faces = [[0, vid(1, k), vid(1, k + 1)] for k in range(n_sect)]
    for r in range(1, n_rings):
        for k in range(n_sect):
            a, b, c, d = vid(r, k), vid(r, k + 1), vid(r + 1, k), vid(r + 1, k + 1)
            faces += [[a, c, d], [a, d, b]] if (r + k) % 2 else [[a, c, b], [b, c, d]]
    return np.array(pts), np.array(faces, np.uint32)


def draw_sources(rng: np.random.Generator, nv: int, k: int) -> np.ndarray:
    """``k`` distinct source vertices drawn uniformly without replacement, in
    draw order (source q gets cluster id q + 1)."""
    return rng.choice(nv, size=k, replace=False).astype(np.uint32)
