"""Survival perceptron: the FP64 numpy reference is pinned by finite
differences on the CPU; the HIP forward, backward and Adam kernels are
compared with it on the GPU (FP32 GEMMs: relative tolerance 2e-4 on the
gradient vector's norm, 1e-5 absolute on probabilities)."""
import numpy as np
import pytest

import mlp_ref


def _params(n_in, hidden, seed):
    rng = np.random.default_rng(seed)
    dims = [n_in, hidden, hidden, hidden, 1]
    ps = []
    for l in range(4):
        lim = np.sqrt(6.0 / dims[l])
        ps.append(rng.uniform(-lim, lim, dims[l] * dims[l + 1]))
        ps.append(rng.normal(0, 0.1, dims[l + 1]))
    return np.concatenate(ps)


def test_reference_gradient_matches_finite_differences():
    n_in, hidden = 4, 7
    p = _params(n_in, hidden, 3)
    rng = np.random.default_rng(4)
    X = rng.normal(size=(23, n_in))
    y = (rng.random(23) < 0.5).astype(float)
    g, l = mlp_ref.gradient(p, X, y, n_in, hidden)
    assert abs(l - mlp_ref.loss(p, X, y, n_in, hidden)) < 1e-14
    h = 1e-6
    fd = np.empty_like(p)
    for i in range(p.size):
        e = np.zeros_like(p)
        e[i] = h
        fd[i] = (mlp_ref.loss(p + e, X, y, n_in, hidden) - mlp_ref.loss(p - e, X, y, n_in, hidden)) / (2 * h)
    assert np.max(np.abs(fd - g)) < 1e-8 * max(1.0, np.max(np.abs(g)))


def test_reference_loss_closed_form():
    # all-zero weights: z = b4, loss = softplus(b4) - mean(y) b4
    n_in, hidden = 3, 5
    p = np.zeros(n_in * hidden + hidden + 2 * (hidden * hidden + hidden) + hidden + 1)
    p[-1] = 0.7
    y = np.array([1.0, 0.0, 1.0, 1.0])
    X = np.ones((4, n_in))
    assert abs(mlp_ref.loss(p, X, y, n_in, hidden) - (np.log1p(np.exp(0.7)) - 0.75 * 0.7)) < 1e-15
    g, _ = mlp_ref.gradient(p, X, y, n_in, hidden)
    assert abs(g[-1] - (1 / (1 + np.exp(-0.7)) - 0.75)) < 1e-15


@pytest.mark.gpu
@pytest.mark.parametrize("n_in,hidden,B", [(19, 256, 4096), (19, 256, 300), (5, 40, 77)])
def test_gradient_matches_reference(tracker, n_in, hidden, B):
    from elegant_amd import learned
    m = learned.MLP(tracker, n_in=n_in, hidden=hidden, seed=11)
    p = _params(n_in, hidden, 5).astype(np.float32)
    m.set_params(p)
    assert np.array_equal(m.params(), p)
    rng = np.random.default_rng(B)
    X = rng.normal(size=(B, n_in)).astype(np.float32)
    y = (rng.random(B) < 0.4).astype(np.float32)
    g, l = m.gradient(X, y)
    gr, lr = mlp_ref.gradient(p.astype(np.float64), X.astype(np.float64), y, n_in, hidden)
    assert abs(l - lr) < 1e-5 * max(1.0, abs(lr))
    assert np.linalg.norm(g - gr) < 2e-4 * np.linalg.norm(gr)
    W = mlp_ref.unpack(gr, n_in, hidden)
    Wg = mlp_ref.unpack(g, n_in, hidden)
    for a, b in zip(W, Wg):  # every layer and bias, not only the bulk
        assert np.linalg.norm(a - b) <= 2e-4 * np.linalg.norm(a) + 1e-9
    prob = m.predict(X)
    assert np.max(np.abs(prob - mlp_ref.predict(p, X, n_in, hidden))) < 1e-5
    m.close()


@pytest.mark.gpu
def test_training_learns_a_survival_boundary(tracker):
    """points survive inside an ellipse whose size shrinks with |delta| and an
    error feature; Adam on the GPU must cut the loss fourfold and classify
    held-out points with >95 % accuracy"""
    from elegant_amd import learned
    rng = np.random.default_rng(0)
    n = 65536
    X = np.zeros((n, 19), dtype=np.float32)
    X[:, 0] = rng.random(n)
    X[:, 1] = rng.random(n)
    X[:, 2] = rng.uniform(-1, 1, n)
    X[:, 3:] = rng.normal(0, 1, (n, 16))
    r = 0.6 - 0.2 * X[:, 2] ** 2 + 0.05 * X[:, 3]
    y = (X[:, 0] ** 2 + X[:, 1] ** 2 < r ** 2).astype(np.float32)
    m = learned.MLP(tracker, seed=2)
    losses = m.train(X[:49152], y[:49152], epochs=24, batch=512, lr=2e-3, seed=9)
    assert losses[-1] < 0.25 * losses[0]
    p = m.predict(X[49152:])
    acc = np.mean((p >= 0.5) == (y[49152:] > 0.5))
    assert acc > 0.95, (acc, losses)
    m.close()
