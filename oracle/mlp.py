"""TEST INFRASTRUCTURE -- NOT PART OF THE PRODUCT.

FP64 numpy restatement of the learned band predictor (no reference
counterpart: the reference has no learned component).  Used only by tests/
to check the library's GPU feature kernel and BF16-MFMA forward/backward
kernels.  Pins: the analytic backward pass is checked against central finite
differences in tests/test_oracle_mlp.py; features against hand values on a
tetrahedron and a plane.

Per-vertex features (10), for vertex v with nearest source s (Euclidean),
e = |p - p_s|, u = (p - p_s)/e, mean fan edge h of the mesh, link vertices of
v (che::link order) L, fan triangles (v, L_k, L_k+1):
  f0 = log1p(e / h)
  f1 = n_v . u                    n_v: normalised sum of fan triangle normals
  f2 = 2 (mean(L) - p) . n_v / lbar    lbar: mean |L_k - p|
  f3 = |L| / 6                    valence
  f4 = level * h / (e + h)        level: topleset index
  f5 = (p - p_s) . n_s / (e + h)  n_s: normal at s
  f6 = lbar / h
  f7 = n_v . n_s
  f8 = level / max level
  f9 = min_k |L_k - p| / max_k |L_k - p|
Target: log(d / e) for e > 0; prediction dhat = e * exp(y).
Network: 10 -> 64 -> 64 -> 1, ReLU, inputs padded to 32.
"""
from __future__ import annotations

import numpy as np

NF = 10
KPAD = 32
H = 64


def fan_lists(mesh):
    """Per vertex: (link vertex list, fan triangle list (a, b)) from the
    oracle CHE (star walk from EVT, include/che.h:10)."""
    links, tris = [], []
    for v in range(mesh.nv):
        st = mesh.star(v)
        L = [int(mesh.vt[(h // 3) * 3 + (h + 1) % 3]) for h in st]
        if len(st):
            last = st[-1]
            b = int(mesh.vt[(last // 3) * 3 + (last + 2) % 3])
            closed = int(mesh.ot[(last // 3) * 3 + (last + 2) % 3]) == int(st[0])
            fan = L + [L[0] if closed else b]
            link = L if closed else L + [b]
        else:
            fan, link = [], []
        links.append(link)
        tris.append([(fan[k], fan[k + 1]) for k in range(len(fan) - 1)])
    return links, tris


def normals(xyz, tris):
    n = np.zeros_like(xyz)
    for v, tl in enumerate(tris):
        acc = np.zeros(3)
        for a, b in tl:
            acc += np.cross(xyz[a] - xyz[v], xyz[b] - xyz[v])
        nn = np.linalg.norm(acc)
        n[v] = acc / nn if nn > 0 else 0
    return n


def features(mesh, xyz, sources, levels, mean_edge):
    xyz = np.asarray(xyz, np.float64)
    links, tris = fan_lists(mesh)
    nrm = normals(xyz, tris)
    src = np.asarray(sources)
    dd = np.linalg.norm(xyz[:, None, :] - xyz[src][None, :, :], axis=2)
    near = src[np.argmin(dd, axis=1)]
    e = dd.min(axis=1)
    h = mean_edge
    lev = levels.astype(np.float64)
    reached = levels != 0xFFFFFFFF
    lmax = max(1.0, lev[reached].max())
    F = np.zeros((len(xyz), NF))
    for v in range(len(xyz)):
        L = links[v]
        p = xyz[v]
        if L:
            ed = np.linalg.norm(xyz[L] - p, axis=1)
            lbar = ed.mean()
            curv = 2 * (xyz[L].mean(axis=0) - p) @ nrm[v] / lbar if lbar > 0 else 0.0
            emin_emax = ed.min() / ed.max() if ed.max() > 0 else 0.0
        else:
            lbar, curv, emin_emax = 0.0, 0.0, 0.0
        u = (p - xyz[near[v]]) / e[v] if e[v] > 0 else np.zeros(3)
        ns = nrm[near[v]]
        lv = lev[v] if reached[v] else 0.0
        F[v] = [np.log1p(e[v] / h), nrm[v] @ u, curv, len(L) / 6.0, lv * h / (e[v] + h),
                (p - xyz[near[v]]) @ ns / (e[v] + h), lbar / h, nrm[v] @ ns, lv / lmax, emin_emax]
    return F, e


def init_params(seed=0):
    rng = np.random.default_rng(seed)
    return {
        "W1": rng.normal(0, np.sqrt(2.0 / NF), (H, KPAD)) * (np.arange(KPAD) < NF),
        "b1": np.zeros(H),
        "W2": rng.normal(0, np.sqrt(2.0 / H), (H, H)),
        "b2": np.zeros(H),
        "w3": rng.normal(0, np.sqrt(1.0 / H), H),
        "b3": np.zeros(1),
    }


def forward(X, p):
    """X: [B, 32].  Returns y [B] and the cache for backward."""
    z1 = X @ p["W1"].T + p["b1"]
    h1 = np.maximum(z1, 0)
    z2 = h1 @ p["W2"].T + p["b2"]
    h2 = np.maximum(z2, 0)
    y = h2 @ p["w3"] + p["b3"][0]
    return y, (X, z1, h1, z2, h2)


def loss(y, t):
    return float(np.mean((y - t) ** 2))


def backward(y, t, cache, p):
    """Gradients of mean((y - t)^2) with respect to every parameter."""
    X, z1, h1, z2, h2 = cache
    B = len(y)
    dy = 2.0 * (y - t) / B
    g = {"w3": h2.T @ dy, "b3": np.array([dy.sum()])}
    dh2 = np.outer(dy, p["w3"]) * (z2 > 0)
    g["W2"] = dh2.T @ h1
    g["b2"] = dh2.sum(axis=0)
    dh1 = (dh2 @ p["W2"]) * (z1 > 0)
    g["W1"] = dh1.T @ X
    g["b1"] = dh1.sum(axis=0)
    return g


def bf16(a):
    """Round to bfloat16 (round-to-nearest-even) and back to float64."""
    a32 = np.asarray(a, np.float32)
    u = a32.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)
