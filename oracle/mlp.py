"""CPU restatement (float64 numpy) of the subtree-cost model of the learned schedule.

ORACLE — test infrastructure only: imported by tests/ as the checker of the HIP forward,
gradient and training kernels (cudqc_amd/csrc/costmodel.hip), never by the product.

The model has no counterpart in the reference (its tasks run in arrival order,
src/host_functions.cu:173-285). Layout of the flat parameter vector: W1[64][7], b1[64],
W2[64][64], b2[64], W3[64], b3. Loss: mean squared error of the scalar output.
"""
import numpy as np

IN, H = 7, 64
SIZES = [("W1", H * IN), ("b1", H), ("W2", H * H), ("b2", H), ("W3", H), ("b3", 1)]
NPARAM = sum(n for _, n in SIZES)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    out, o = {}, 0
    for name, n in SIZES:
        out[name] = p[o:o + n]
        o += n
    out["W1"] = out["W1"].reshape(H, IN)
    out["W2"] = out["W2"].reshape(H, H)
    return out


def features(nm, tot, depth, dmin_o, dmax_o, dmin_i, dmax_i):
    """log1p of members, candidates, depth and the four degree extremes."""
    return np.log1p(np.stack([nm, np.asarray(tot) - np.asarray(nm), depth, dmin_o, dmax_o, dmin_i, dmax_i], axis=-1)
                    .astype(np.float64))


def forward(p, x):
    w = unpack(p)
    x = np.asarray(x, dtype=np.float64)
    z1 = x @ w["W1"].T + w["b1"]
    h1 = np.maximum(z1, 0.0)
    z2 = h1 @ w["W2"].T + w["b2"]
    h2 = np.maximum(z2, 0.0)
    return h2 @ w["W3"] + w["b3"][0]


def loss_and_grad(p, x, t):
    """Mean squared error and its gradient with respect to the flat parameters."""
    w = unpack(p)
    x = np.asarray(x, dtype=np.float64)
    t = np.asarray(t, dtype=np.float64)
    n = len(x)
    z1 = x @ w["W1"].T + w["b1"]
    h1 = np.maximum(z1, 0.0)
    z2 = h1 @ w["W2"].T + w["b2"]
    h2 = np.maximum(z2, 0.0)
    y = h2 @ w["W3"] + w["b3"][0]
    r = y - t
    loss = float(np.mean(r * r))
    dy = 2.0 * r / n
    g = {"W3": h2.T @ dy, "b3": np.array([dy.sum()])}
    dz2 = np.outer(dy, w["W3"]) * (z2 > 0)
    g["W2"] = dz2.T @ h1
    g["b2"] = dz2.sum(axis=0)
    dz1 = (dz2 @ w["W2"]) * (z1 > 0)
    g["W1"] = dz1.T @ x
    g["b1"] = dz1.sum(axis=0)
    return loss, np.concatenate([g[name].ravel() for name, _ in SIZES])
