"""Trains the learned band predictor on the GPU.

Data: `n_meshes` seeded meshes (height fields of 8 random sinusoids with a
random amplitude scale, and displaced icospheres), vertex counts log-uniform
in [10k, 300k]; one random source per mesh.  Labels are GPU distance fields
from the patch schedule run to tol 1e-6 (the fixed point of the
relaxation); features are computed on the GPU
(gp_mlp_features); `samples_per_mesh` vertices are drawn per mesh.  The MLP
is trained with the library's BF16-MFMA gradient kernel and Adam
(gp_mlp_train).  The parameters are saved as a float32 .npy.

    python -m gproshan_amd.train --meshes 2000 --out gproshan_amd/models/band_mlp.npy
"""
from __future__ import annotations

import argparse
import math
import os
import sys
import time

import numpy as np

from . import _lib as L
from . import geodesics as P
from . import learned as Lr
from . import meshgen as G

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_MODEL = os.path.join(HERE, "models", "band_mlp.npy")


def make_mesh(rng: np.random.Generator):
    """One training mesh: ~70% height fields, ~30% displaced icospheres."""
    nv = math.exp(rng.uniform(math.log(10_000), math.log(300_000)))
    if rng.random() < 0.7:
        side = max(64, int(round(math.sqrt(nv))))
        rows = max(64, int(side * rng.uniform(0.7, 1.3)))
        cols = max(64, int(nv / rows))
        xyz, f = G.height_field_rect(rows, cols, seed=int(rng.integers(1 << 31)))
        xyz[:, 2] *= rng.uniform(0.25, 2.0)           # vary the relief
    else:
        level = int(np.clip(round(math.log(nv / 10.0) / math.log(4.0)), 5, 7))
        xyz, f = G.icosphere(level, float(rng.uniform(0.0, 0.02)), seed=int(rng.integers(1 << 31)))
    return xyz, f


def label_field(mesh, src):
    plan = P.Plan(mesh, src, schedule=L.SCHED_PATCHES, tol=1e-6)
    plan.run()
    return plan.result()


def build_dataset(n_meshes: int, samples_per_mesh: int, seed: int, log=print):
    rng = np.random.default_rng(seed)
    Fs, ts = [], []
    t0 = time.time()
    for i in range(n_meshes):
        xyz, f = make_mesh(rng)
        mesh = P.che(xyz, f)
        src = G.draw_sources(rng, len(xyz), 1)
        d = label_field(mesh, src)
        F, e = Lr.features(mesh, src)
        ok = np.flatnonzero(np.isfinite(d) & (e > 0) & (d > 0))
        pick = rng.choice(ok, size=min(samples_per_mesh, len(ok)), replace=False)
        Fs.append(F[pick])
        ts.append(np.log(d[pick] / e[pick]).astype(np.float32))
        if (i + 1) % 50 == 0 or i + 1 == n_meshes:
            log(f"[train] data {i + 1}/{n_meshes} meshes, {sum(len(x) for x in ts)} samples, {time.time() - t0:.0f}s")
    return np.concatenate(Fs), np.concatenate(ts)


def train(n_meshes=2000, samples_per_mesh=2048, steps=3000, batch=16384, lr=2e-3, seed=0, out=DEFAULT_MODEL,
          log=print):
    F, t = build_dataset(n_meshes, samples_per_mesh, seed, log)
    net = Lr.MLP(seed=seed)
    _, l0 = net.gradient(F[:65536], t[:65536])
    t0 = time.time()
    done = 0
    while done < steps:
        k = min(500, steps - done)
        last = net.train(F, t, steps=k, batch=batch, lr=lr, seed=seed + done)
        done += k
        log(f"[train] step {done}/{steps} batch loss {last:.4e} ({time.time() - t0:.0f}s)")
    _, l1 = net.gradient(F[:65536], t[:65536])
    log(f"[train] loss {l0:.4e} -> {l1:.4e} on {len(F)} samples; target variance {np.var(t):.4e}")
    p = net.params()
    if out:
        os.makedirs(os.path.dirname(out), exist_ok=True)
        np.save(out, p)
        log(f"[train] saved {out}")
    return net, dict(loss0=l0, loss1=l1, samples=len(F), target_var=float(np.var(t)))


def load(path=DEFAULT_MODEL):
    p = np.load(path, allow_pickle=False)
    return Lr.MLP(params=p)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--meshes", type=int, default=2000)
    ap.add_argument("--samples", type=int, default=2048)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--batch", type=int, default=16384)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", default=DEFAULT_MODEL)
    a = ap.parse_args(argv)
    train(a.meshes, a.samples, a.steps, a.batch, a.lr, a.seed, a.out, log=lambda *x: print(*x, flush=True))


if __name__ == "__main__":
    sys.exit(main())
