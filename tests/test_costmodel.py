"""The subtree-cost model of the learned schedule.

CPU: the float64 restatement (oracle/mlp.py) is pinned by central finite differences and
by closed forms of tiny cases; the library's host feature transform equals it.
GPU: the HIP forward, gradient and training kernels against that restatement.
"""
import numpy as np
import pytest

from oracle import mlp as omlp


def random_params(seed, scale=0.3):
    rng = np.random.default_rng(seed)
    return rng.normal(0, scale, omlp.NPARAM)


def random_data(seed, n):
    rng = np.random.default_rng(seed)
    x = rng.uniform(0, 4, (n, omlp.IN))
    t = rng.normal(3, 1, n)
    return x, t


def test_oracle_gradient_matches_finite_differences():
    p = random_params(1)
    x, t = random_data(2, 37)
    _, g = omlp.loss_and_grad(p, x, t)
    rng = np.random.default_rng(3)
    idx = np.concatenate([rng.choice(omlp.NPARAM, 60, replace=False), [omlp.NPARAM - 1, 0, 448, 4672]])
    eps = 1e-6
    for i in idx:
        d = np.zeros_like(p)
        d[i] = eps
        num = (omlp.loss_and_grad(p + d, x, t)[0] - omlp.loss_and_grad(p - d, x, t)[0]) / (2 * eps)
        assert abs(num - g[i]) <= 1e-6 + 1e-5 * abs(num), (i, num, g[i])


def test_oracle_forward_closed_form():
    """With W2 = identity, b = 0 and W3 = ones, the output is the sum of relu(W1 x)."""
    w = np.zeros(omlp.NPARAM)
    rng = np.random.default_rng(4)
    W1 = rng.normal(size=(64, 7))
    w[:448] = W1.ravel()
    w[512:4608] = np.eye(64).ravel()
    w[4672:4736] = 1.0
    x = rng.uniform(-1, 1, (5, 7))
    assert np.allclose(omlp.forward(w, x), np.maximum(x @ W1.T, 0).sum(axis=1))


def test_host_features_match_restatement(dqc):
    recs = [dqc.TaskLog(uid=i, parent=0, depth=d, nm=nm, tot=tot, dmin_out=a, dmax_out=b, dmin_in=c, dmax_in=e,
                        branches=0, flags=0)
            for i, (d, nm, tot, a, b, c, e) in enumerate([(0, 0, 13452, 16, 900, 16, 850), (7, 12, 40, 11, 39, 12, 38),
                                                          (65535, 3, 3, 0, 2, 0, 2)])]
    x = dqc.task_features(recs)
    ref = omlp.features(*[np.array([getattr(r, f) for r in recs]) for f in
                          ("nm", "tot", "depth", "dmin_out", "dmax_out", "dmin_in", "dmax_in")])
    assert np.allclose(x, ref, rtol=1e-6)


@pytest.mark.gpu
def test_gpu_forward_matches_oracle(dqc):
    p = random_params(5)
    x, _ = random_data(6, 3001)
    y = dqc.mlp_forward(p, x)
    ref = omlp.forward(p.astype(np.float32).astype(np.float64), x.astype(np.float32).astype(np.float64))
    assert np.allclose(y, ref, rtol=1e-4, atol=1e-4)


@pytest.mark.gpu
def test_gpu_gradient_matches_oracle(dqc):
    """HIP gradient vs the float64 analytic gradient: relative error <= 1e-3 overall and
    per parameter group."""
    p = random_params(7).astype(np.float32)
    x, t = random_data(8, 2500)
    x, t = x.astype(np.float32), t.astype(np.float32)
    g, loss = dqc.mlp_gradient(p, x, t)
    rloss, rg = omlp.loss_and_grad(p.astype(np.float64), x.astype(np.float64), t.astype(np.float64))
    assert abs(loss - rloss) <= 1e-4 * abs(rloss)
    o = 0
    for name, n in omlp.SIZES:
        a, b = g[o:o + n], rg[o:o + n]
        assert np.linalg.norm(a - b) <= 1e-3 * np.linalg.norm(b) + 1e-7, name
        o += n


@pytest.mark.gpu
def test_gpu_training_fits_a_known_function(dqc):
    """Adam on the GPU drives the loss of a smooth target far below its variance."""
    rng = np.random.default_rng(9)
    x = rng.uniform(0, 3, (8192, 7)).astype(np.float32)
    t = (np.sin(x[:, 0]) + 0.5 * x[:, 1] * x[:, 2] - 0.3 * x[:, 4]).astype(np.float32)
    p0 = dqc.mlp_init(1)
    p, loss = dqc.mlp_train(p0, x, t, epochs=60, batch=256, lr=3e-3, seed=2)
    assert loss < 0.02 * float(np.var(t)), loss
    ref = omlp.loss_and_grad(p.astype(np.float64), x.astype(np.float64), t.astype(np.float64))[0]
    assert abs(ref - loss) <= 1e-3 * ref + 1e-6
