"""Pins of the CPU reference of the surrogate network: its hand-written
backward against torch autograd (float64) and central finite differences on
a tiny random graph.  torch is imported inside the tests: a module-level
import would load torch's own HIP runtime into every test process at
collection, GPU runs included."""
import numpy as np
import pytest

import gnn_ref as G


def tiny_graph(n=40, seed=0):
    rng = np.random.default_rng(seed)
    raw = np.zeros((n, G.RAW), np.float32)
    raw[:, 0:2] = rng.uniform(0, 0.01, (n, 2))
    raw[:, 2:7] = rng.normal(0, 1, (n, 5))
    raw[:, 7] = 0.003
    raw[:, 8] = rng.uniform(0, 1, n) < 0.2
    nb = -np.ones((n, G.DEG), np.int32)
    nd = np.zeros(n, np.int32)
    for i in range(n):
        d = np.hypot(raw[:, 0] - raw[i, 0], raw[:, 1] - raw[i, 1])
        js = [j for j in np.argsort(d) if j != i and d[j] < 0.006][:G.DEG]
        nb[i, :len(js)] = js
        nd[i] = len(js)
    nf = rng.normal(0, 1, (n, G.FN)).astype(np.float32)
    nf[:, 0] = 1.0
    tgt = rng.normal(0, 1, (n, G.NO))
    tgt[:, 3] = 0.0
    tw = (rng.uniform(0, 1, n) < 0.8).astype(np.float64)
    w = rng.normal(0, 0.3, G.NPARAM)
    w[G.O_W5:G.O_B5].reshape(G.D, G.NOP)[:, G.NO:] = 0.0
    w[G.O_B5 + G.NO:] = 0.0
    return w, nf, raw, nb, nd, tgt, tw


def torch_loss(w, nf, e, ii, n, tgt, tw):
    import torch
    p = {k: torch.tensor(v, dtype=torch.float64) for k, v in G.unpack(w).items()}
    for v in p.values():
        v.requires_grad_(True)
    E = torch.tensor(e, dtype=torch.float64)
    a1 = torch.relu(E @ p["W1"])
    hs = torch.zeros(n, G.HE, dtype=torch.float64).index_add(0, torch.tensor(ii), a1) * G.EDGE_NORM
    m = hs @ p["W2"]
    x0 = torch.cat([torch.tensor(nf, dtype=torch.float64), m], 1)
    h3 = torch.relu(x0 @ p["W3"] + p["b3"])
    h4 = torch.relu(h3 @ p["W4"] + p["b4"])
    y = (h4 @ p["W5"] + p["b5"])[:, :G.NO]
    L = (torch.tensor(tw)[:, None] * (y - torch.tensor(tgt)) ** 2).sum()
    L.backward()
    g = {k: v.grad.numpy() for k, v in p.items()}
    g["W5"][:, G.NO:] = 0.0
    return float(L.detach()), G.pack(g)


def test_backward_matches_autograd():
    w, nf, raw, nb, nd, tgt, tw = tiny_graph()
    ii, jj = G.edges(nb, nd)
    e = G.edge_features(raw, ii, jj)
    L, g, _ = G.loss_and_grad(w, nf, e, ii, len(nd), tgt, tw)
    Lt, gt = torch_loss(w, nf, e, ii, len(nd), tgt, tw)
    assert abs(L - Lt) <= 1e-12 * abs(Lt)
    g[G.O_B5 + G.NO:] = 0.0
    gt[G.O_B5 + G.NO:] = 0.0
    np.testing.assert_allclose(g, gt, rtol=1e-10, atol=1e-12 * np.abs(gt).max())


@pytest.mark.parametrize("block", ["W1", "W2", "W3", "b3", "W4", "b4", "W5", "b5"])
def test_backward_matches_finite_differences(block):
    w, nf, raw, nb, nd, tgt, tw = tiny_graph(seed=1)
    ii, jj = G.edges(nb, nd)
    e = G.edge_features(raw, ii, jj)
    _, g, _ = G.loss_and_grad(w, nf, e, ii, len(nd), tgt, tw)
    off = {"W1": G.O_W1, "W2": G.O_W2, "W3": G.O_W3, "b3": G.O_B3, "W4": G.O_W4, "b4": G.O_B4,
           "W5": G.O_W5, "b5": G.O_B5}[block]
    rng = np.random.default_rng(5)
    size = {"W1": G.FE * G.HE, "W2": G.HE * G.HE, "W3": G.FI * G.D, "b3": G.D, "W4": G.D * G.D,
            "b4": G.D, "W5": G.D * G.NOP, "b5": G.NO}[block]
    idx = off + rng.choice(size, min(12, size), replace=False)
    if block == "W5":
        idx = np.array([G.O_W5 + r * G.NOP + c for r, c in zip(rng.integers(0, G.D, 12), rng.integers(0, G.NO, 12))])
    h = 1e-6
    for k in idx:
        wp, wm = w.copy(), w.copy()
        wp[k] += h
        wm[k] -= h
        Lp = G.loss_and_grad(wp, nf, e, ii, len(nd), tgt, tw)[0]
        Lm = G.loss_and_grad(wm, nf, e, ii, len(nd), tgt, tw)[0]
        fd = (Lp - Lm) / (2 * h)
        assert abs(fd - g[k]) <= 1e-5 * max(1.0, abs(fd)), (block, k, fd, g[k])


def test_bf16_rounding_is_nearest_even():
    # bf16 keeps 7 mantissa bits: the spacing at 1.0 is 2^-7
    x = np.array([1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, 1.0 + 2 ** -9, 1.0 + 3 * 2 ** -8, -2.5], np.float32)
    r = G.bf16(x)
    assert r[0] == 1.0                  # tie: to the even mantissa
    assert r[1] == 1.0 + 2 ** -7        # above the tie
    assert r[2] == 1.0                  # below the tie
    assert r[3] == 1.0 + 2 ** -6        # tie between odd 1+2^-7 and even 1+2^-6
    assert r[4] == -2.5
