"""CPU reference of the learned surrogate's network (test infrastructure).

Restates, in float64 numpy, the network of sources_amd/csrc/sph_gnn.h:
    z_ij = e_ij W1;  m_i = (sum_j relu(z_ij) / 16) W2
    y_i  = W5 relu(W4 relu(W3 [n_i, m_i] + b3) + b4) + b5
with a hand-written backward.  `rb=True` rounds every matrix-product operand
to BF16 (round to nearest even) at the points where the MFMA kernels do, so
that the GPU's forward and gradients can be compared; `rb=False` is the
exact float64 network, which autograd and finite differences pin.
"""
import numpy as np

FE, HE, FN, D, NO, NOP, DEG, RAW = 16, 32, 32, 64, 4, 32, 48, 12
FI = FN + HE
EDGE_NORM = 1.0 / 16.0
O_W1 = 0
O_W2 = O_W1 + FE * HE
O_W3 = O_W2 + HE * HE
O_B3 = O_W3 + FI * D
O_W4 = O_B3 + D
O_B4 = O_W4 + D * D
O_W5 = O_B4 + D
O_B5 = O_W5 + D * NOP
NPARAM = O_B5 + NOP


def bf16(x):
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    x = np.asarray(x, np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(w):
    w = np.asarray(w, np.float64)
    return {
        "W1": w[O_W1:O_W2].reshape(FE, HE), "W2": w[O_W2:O_W3].reshape(HE, HE),
        "W3": w[O_W3:O_B3].reshape(FI, D), "b3": w[O_B3:O_W4],
        "W4": w[O_W4:O_B4].reshape(D, D), "b4": w[O_B4:O_W5],
        "W5": w[O_W5:O_B5].reshape(D, NOP), "b5": w[O_B5:NPARAM],
    }


def pack(p):
    out = np.zeros(NPARAM)
    for k, off, end in (("W1", O_W1, O_W2), ("W2", O_W2, O_W3), ("W3", O_W3, O_B3), ("b3", O_B3, O_W4),
                        ("W4", O_W4, O_B4), ("b4", O_B4, O_W5), ("W5", O_W5, O_B5), ("b5", O_B5, NPARAM)):
        out[off:end] = np.asarray(p[k]).ravel()
    return out


def edges(nb, nd):
    """Edge list (i, j, slot) in the kernel's slot order."""
    n = len(nd)
    ii, ss = np.nonzero(np.arange(DEG)[None, :] < nd[:, None])
    jj = nb[ii, ss]
    assert (jj >= 0).all() and (jj < n).all()
    return ii, jj


def edge_features(raw, ii, jj):
    """The kernel's edge features, computed in float32 like the kernel."""
    raw = np.asarray(raw, np.float32)
    ri, rj = raw[ii], raw[jj]
    ih = np.float32(1.0) / ri[:, 7]
    dx = ri[:, 0] - rj[:, 0]
    dy = ri[:, 1] - rj[:, 1]
    r = np.sqrt(dx * dx + dy * dy)
    q = np.minimum(r * ih * np.float32(0.5), np.float32(1.0))
    e = np.zeros((len(ii), FE), np.float32)
    e[:, 0] = 1.0
    e[:, 1] = dx * ih
    e[:, 2] = dy * ih
    e[:, 3] = r * ih
    e[:, 4] = (np.float32(1.0) - q) * (np.float32(1.0) - q)
    e[:, 5] = ri[:, 2] - rj[:, 2]
    e[:, 6] = ri[:, 3] - rj[:, 3]
    e[:, 7] = ri[:, 4] - rj[:, 4]
    e[:, 8] = ri[:, 5] - rj[:, 5]
    e[:, 9] = ri[:, 6] - rj[:, 6]
    e[:, 10] = rj[:, 8]
    return e


def forward(w, nf, e, ii, n, rb=False):
    R = bf16 if rb else (lambda a: np.asarray(a, np.float64))
    p = unpack(w)
    E = R(e)
    z1 = E @ R(p["W1"])
    a1 = np.maximum(z1, 0.0)
    hs = np.zeros((n, HE))
    np.add.at(hs, ii, a1)
    hs *= EDGE_NORM
    if rb:
        hs = hs.astype(np.float32).astype(np.float64)
    m = R(hs) @ R(p["W2"])
    x0 = np.concatenate([np.asarray(nf, np.float64), m], axis=1)
    if rb:
        x0 = x0.astype(np.float32).astype(np.float64)
    z3 = R(x0) @ R(p["W3"]) + p["b3"]
    h3 = np.maximum(z3, 0.0)
    z4 = R(h3) @ R(p["W4"]) + p["b4"]
    h4 = np.maximum(z4, 0.0)
    y = R(h4) @ R(p["W5"]) + p["b5"]
    cache = dict(E=E, z1=z1, hs=hs, m=m, x0=x0, h3=h3, h4=h4, ii=ii)
    return y[:, :NO], cache


def loss_and_grad(w, nf, e, ii, n, tgt_n, tw, rb=False):
    """L = sum_i w_i sum_o (y_io - t_io)^2 with t the normalised targets;
    returns L, dL/dW (NPARAM) and y."""
    R = bf16 if rb else (lambda a: np.asarray(a, np.float64))
    p = unpack(w)
    y, c = forward(w, nf, e, ii, n, rb)
    err = y - tgt_n
    L = float(np.sum(tw[:, None] * err * err))
    dy = np.zeros((n, NOP))
    dy[:, :NO] = 2.0 * tw[:, None] * err
    g = {}
    g["W5"] = R(c["h4"]).T @ R(dy)
    g["W5"][:, NO:] = 0.0
    g["b5"] = dy.sum(0)
    dz4 = (R(dy) @ R(p["W5"]).T) * (c["h4"] > 0)
    g["W4"] = R(c["h3"]).T @ R(dz4)
    g["b4"] = dz4.sum(0)
    dz3 = (R(dz4) @ R(p["W4"]).T) * (c["h3"] > 0)
    g["W3"] = R(c["x0"]).T @ R(dz3)
    g["b3"] = dz3.sum(0)
    dm = R(dz3) @ R(p["W3"][FN:]).T
    g["W2"] = R(c["hs"]).T @ R(dm)
    dhs = (R(dm) @ R(p["W2"]).T) * EDGE_NORM
    dz1 = dhs[c["ii"]] * (c["z1"] > 0)
    g["W1"] = c["E"].T @ R(dz1)
    return L, pack(g), y
