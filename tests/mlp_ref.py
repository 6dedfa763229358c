"""FP64 numpy reference of the survival perceptron (test infrastructure).

Forward: h1 = relu(X W1^T + b1), h2 = relu(h1 W2^T + b2),
h3 = relu(h2 W3^T + b3), z = h3 W4^T + b4; loss = mean(softplus(z) - y z).
The backward pass is written out by hand and is itself checked against
central finite differences in tests/test_mlp.py.
"""
import numpy as np


def unpack(params, n_in=19, hidden=256):
    dims = [n_in, hidden, hidden, hidden, 1]
    out = []
    off = 0
    p = np.asarray(params, dtype=np.float64)
    for l in range(4):
        w = p[off:off + dims[l] * dims[l + 1]].reshape(dims[l + 1], dims[l])
        off += w.size
        b = p[off:off + dims[l + 1]]
        off += b.size
        out += [w, b]
    assert off == p.size
    return out


def forward(params, X, n_in=19, hidden=256):
    W = unpack(params, n_in, hidden)
    a = [np.asarray(X, dtype=np.float64)]
    for l in range(4):
        z = a[-1] @ W[2 * l].T + W[2 * l + 1]
        a.append(np.maximum(z, 0.0) if l < 3 else z)
    return a


def loss(params, X, y, n_in=19, hidden=256):
    z = forward(params, X, n_in, hidden)[-1][:, 0]
    y = np.asarray(y, dtype=np.float64)
    return float(np.mean(np.logaddexp(0.0, z) - y * z))


def gradient(params, X, y, n_in=19, hidden=256):
    W = unpack(params, n_in, hidden)
    a = forward(params, X, n_in, hidden)
    B = a[0].shape[0]
    z = a[-1][:, 0]
    y = np.asarray(y, dtype=np.float64)
    dz = ((1.0 / (1.0 + np.exp(-z))) - y)[:, None] / B
    grads = [None] * 8
    for l in range(3, -1, -1):
        grads[2 * l] = dz.T @ a[l]
        grads[2 * l + 1] = dz.sum(axis=0)
        if l > 0:
            dz = (dz @ W[2 * l]) * (a[l] > 0)
    return np.concatenate([g.ravel() for g in grads]), float(np.mean(np.logaddexp(0.0, z) - y * z))


def predict(params, X, n_in=19, hidden=256):
    z = forward(params, X, n_in, hidden)[-1][:, 0]
    return 1.0 / (1.0 + np.exp(-z))
