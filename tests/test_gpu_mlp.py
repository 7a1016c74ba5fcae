"""GPU learned-band predictor: features, BF16-MFMA forward and gradients
against the FP64 numpy oracle (oracle/mlp.py)."""
import numpy as np
import pytest

import mlp as Mo
from gproshan_amd import geodesics as P
from gproshan_amd import learned as Lr
from gproshan_amd import meshgen as G

pytestmark = pytest.mark.gpu


def _data(n=4096, seed=0):
    rng = np.random.default_rng(seed)
    F = rng.normal(size=(n, Lr.NFEAT)).astype(np.float32)
    t = (0.1 * np.tanh(F[:, 0]) + 0.05 * F[:, 1] ** 2).astype(np.float32)
    return F, t


def _holed(n=40, seed=9):
    xyz, f = G.height_field(n, seed)
    r, c = f[:, 0] // n, f[:, 0] % n
    keep = f[~((r >= 10) & (r < 16) & (c >= 12) & (c < 20))]
    used = np.unique(keep)
    remap = np.full(len(xyz), -1, np.int64)
    remap[used] = np.arange(len(used))
    return xyz[used], remap[keep].astype(np.uint32)


@pytest.mark.parametrize("mesh", ["hf33", "holed40", "ico3"])
def test_features_match_oracle(oracle, mesh):
    """All ten per-vertex features: on an open height field, on one with an
    internal hole (border vertices inside the mesh) and on a closed sphere."""
    if mesh == "hf33":
        (xyz, f), src = G.height_field(33, 4), [100, 700]
    elif mesh == "holed40":
        (xyz, f), src = _holed(), [5, 900]
    else:
        (xyz, f), src = G.icosphere(3, 0.01, 4), [0, 300]
    M = oracle.Mesh(xyz, f)
    C = P.che(xyz, f)
    lev = M.toplesets(src)[0]
    # mean fan edge length as the library computes it (closing entries counted twice)
    Fo, eo = Mo.features(M, xyz, src, lev, _mean_edge(M, xyz))
    Fg, eg = Lr.features(C, src)
    assert np.allclose(eg, eo, rtol=1e-6, atol=1e-7)
    assert np.allclose(Fg, Fo, rtol=1e-4, atol=1e-5), np.abs(Fg - Fo).max(axis=0)


def _mean_edge(M, xyz):
    tot, cnt = 0.0, 0
    links, tris = Mo.fan_lists(M)
    for v, tl in enumerate(tris):
        if not tl:
            continue
        fan = [tl[0][0]] + [b for _, b in tl]
        for u in fan:
            tot += np.linalg.norm(xyz[u] - xyz[v])
            cnt += 1
    return tot / cnt


def test_forward_matches_oracle():
    F, _ = _data(1000, 1)
    net = Lr.MLP(seed=5)
    p = Lr.unpack(net.params())
    pb = {k: (Mo.bf16(v) if k in ("W1", "W2") else v) for k, v in p.items()}
    X = np.zeros((len(F), Mo.KPAD))
    X[:, :Mo.NF] = Mo.bf16(F)
    yo, cache = Mo.forward(X, pb)
    yg = net.forward(F)
    # BF16 rounding of the hidden activations: ~1e-2 relative
    err = np.abs(yg - yo) / (np.abs(yo).max() + 1e-6)
    assert err.max() < 3e-2, err.max()


def test_gradient_matches_oracle():
    F, t = _data(2048, 2)
    net = Lr.MLP(seed=7)
    p = Lr.unpack(net.params())
    pb = {k: (Mo.bf16(v) if k in ("W1", "W2") else v) for k, v in p.items()}
    X = np.zeros((len(F), Mo.KPAD))
    X[:, :Mo.NF] = Mo.bf16(F)
    y, cache = Mo.forward(X, pb)
    go = Mo.backward(y, t, cache, pb)
    gg, loss = net.gradient(F, t)
    gd = Lr.unpack(gg)
    assert abs(loss - Mo.loss(y, t)) <= 3e-2 * Mo.loss(y, t) + 1e-6
    for k in ("W1", "b1", "W2", "b2", "w3", "b3"):
        a, b = gd[k].reshape(-1), go[k].reshape(-1)
        rel = np.linalg.norm(a - b) / (np.linalg.norm(b) + 1e-12)
        print(f"[grad] {k} rel {rel:.3e}", flush=True)
        assert rel < 5e-2, (k, rel)


def test_training_reduces_loss():
    F, t = _data(32768, 3)
    net = Lr.MLP(seed=1)
    _, l0 = net.gradient(F, t)
    net.train(F, t, steps=300, batch=4096, lr=3e-3)
    _, l1 = net.gradient(F, t)
    print(f"[train] loss {l0:.4e} -> {l1:.4e}", flush=True)
    assert l1 < 0.2 * l0
