"""The benchmark configurations as seeded synthetic meshes, shared by
bench.py and the full-size tests.

  * c2:  1024 x 1024 grid height field (1,048,576 vertices), one source or
         16 sources drawn from the seed;
  * c3:  256 meshes (grid height fields and displaced icospheres) with
         log-uniform vertex counts in [2k, 200k] and U{1..64} sources each,
         every (mesh, source) pair its own one-source field;
  * c4:  held-out 1024 x 1024 height fields (seeds the trainer never draws)
         for the learned-band schedule;
  * c5:  2048 x 2048 grid stretched 1:100 in x, jitter U(-0.4, 0.4) cells,
         4 sources.
"""
from __future__ import annotations

import math

import numpy as np

from . import meshgen as G

# Seeds of the held-out 1M-vertex height fields.  gproshan_amd/train.py draws
# its meshes' seeds at random from [0, 2^31) and builds meshes of 10k-300k
# vertices, so none of its meshes is one of these.
C4_SEEDS = (9_000_101, 9_000_202)


def c2(seed: int = 0, n_sources: int = 1, n: int = 1024):
    """(xyz, faces, sources) of the height-field workload."""
    xyz, f = G.height_field(n, seed)
    rng = np.random.default_rng(seed + (1000 if n_sources == 1 else 2000))
    return xyz, f, G.draw_sources(rng, len(xyz), n_sources)


def c5(seed: int = 0, n: int = 2048):
    """(xyz, faces, sources) of the precision-stress grid."""
    xyz, f = G.anisotropic_grid(n, 100.0, 0.4, seed)
    return xyz, f, G.draw_sources(np.random.default_rng(seed + 1000), len(xyz), 4)


def c4(k: int = 0, n: int = 1024):
    """(xyz, faces, sources) of held-out height field k: one source."""
    seed = C4_SEEDS[k]
    xyz, f = G.height_field(n, seed)
    return xyz, f, G.draw_sources(np.random.default_rng(seed), len(xyz), 1)


def c3(seed: int = 0, n_meshes: int = 256):
    """[(xyz, faces)] and jobs [(mesh index, source)] of the batched case."""
    rng = np.random.default_rng(seed)
    raw, jobs = [], []
    for k in range(n_meshes):
        nv = math.exp(rng.uniform(math.log(2000), math.log(200000)))
        if rng.random() < 0.75:
            rows = max(16, int(round(math.sqrt(nv) * rng.uniform(0.8, 1.25))))
            cols = max(16, int(round(nv / rows)))
            xyz, f = G.height_field_rect(rows, cols, seed=int(rng.integers(1 << 31)))
        else:
            level = int(min(range(4, 8), key=lambda l: abs(math.log(10 * 4 ** l + 2) - math.log(nv))))
            xyz, f = G.icosphere(level, 0.01, seed=int(rng.integers(1 << 31)))
        raw.append((xyz, f))
        for s in G.draw_sources(rng, len(xyz), int(rng.integers(1, 65))):
            jobs.append((k, int(s)))
    return raw, jobs
