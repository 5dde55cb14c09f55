"""numpy (float64) restatement of the row-count predictor.

TEST INFRASTRUCTURE ONLY: imported by tests/ to check the predictor's HIP
inference and training kernels, never by the product package.  The
predictor has no reference counterpart (the reference sizes nothing); its
definition is the one in include/spgemm.h:

  features of C row i (A m x K, B K x n):
    r_i   = nnz(A row i)
    ub_i  = sum over A's entries (i, k) of b_k,  b_k = nnz(B row k)
    max_i = max of those b_k,  sd_i = population standard deviation of them
    x_i   = log1p([r_i, ub_i, max_i, sd_i])
  network: h1 = relu(W1 z + b1), h2 = relu(W2 h1 + b2), y = W3 . h2 + b3,
    z = (x - mu) / sd, parameters packed as W1 (64x4, row major), b1, W2
    (64x64), b2, W3 (64), b3
  baseline: s_i = n * (1 - exp(-ub_i / n)), the distinct count of ub_i
    products falling uniformly on n columns
  target: t_i = log1p(nnz(C row i)) - log1p(s_i)
  prediction: clamp(expm1(y + log1p(s_i)), 0, min(ub_i, n))
  loss: mean over samples of (y - t)^2

Parity pins (tests/test_oracle.py): the gradient against central finite
differences of the loss; the features against a per-row Python loop.
"""
import numpy as np

F, H = 4, 64
NP = H * F + H + H * H + H + H + 1


def unpack(w):
    w = np.asarray(w, np.float64)
    o = 0
    W1 = w[o:o + H * F].reshape(H, F); o += H * F
    b1 = w[o:o + H]; o += H
    W2 = w[o:o + H * H].reshape(H, H); o += H * H
    b2 = w[o:o + H]; o += H
    W3 = w[o:o + H]; o += H
    b3 = w[o]
    return W1, b1, W2, b2, W3, b3


def pack(W1, b1, W2, b2, W3, b3):
    return np.concatenate([W1.ravel(), b1, W2.ravel(), b2, W3, [b3]])


def features(m, a_rows, a_cols, b_ptr):
    """Row features from A's entries (a_rows, a_cols, duplicates counted as
    separate entries) and B's CSR row pointer."""
    a_rows = np.asarray(a_rows, np.int64)
    blen = np.diff(np.asarray(b_ptr, np.int64))
    b = blen[np.asarray(a_cols, np.int64)].astype(np.float64)
    r = np.bincount(a_rows, minlength=m).astype(np.float64)
    s = np.bincount(a_rows, weights=b, minlength=m)
    s2 = np.bincount(a_rows, weights=b * b, minlength=m)
    mx = np.zeros(m)
    np.maximum.at(mx, a_rows, b)
    with np.errstate(invalid="ignore", divide="ignore"):
        mean = np.where(r > 0, s / np.maximum(r, 1), 0.0)
        var = np.where(r > 0, s2 / np.maximum(r, 1) - mean * mean, 0.0)
    var = np.maximum(var, 0.0)
    return np.log1p(np.stack([r, s, mx, np.sqrt(var)], axis=1)), s.astype(np.int64)


def forward(w, norm, x):
    W1, b1, W2, b2, W3, b3 = unpack(w)
    z = (np.asarray(x, np.float64) - norm[:4]) / norm[4:]
    h1 = np.maximum(z @ W1.T + b1, 0.0)
    h2 = np.maximum(h1 @ W2.T + b2, 0.0)
    return h2 @ W3 + b3


def log1p_uniform_count(ub, n):
    ub = np.asarray(ub, np.float64)
    return np.log1p(n * -np.expm1(-ub / n))


def target(cnt, ub, n):
    return np.log1p(np.asarray(cnt, np.float64)) - log1p_uniform_count(ub, n)


def predict(w, norm, x, ub, n):
    y = forward(w, norm, x)
    ub = np.asarray(ub, np.float64)
    p = np.expm1(y + log1p_uniform_count(ub, n))
    p = np.clip(p, 0.0, np.minimum(ub, n))
    return np.where(ub > 0, p, 0.0)


def loss_and_grad(w, norm, x, t):
    """Mean squared error and its gradient by backpropagation."""
    W1, b1, W2, b2, W3, b3 = unpack(w)
    z = (np.asarray(x, np.float64) - norm[:4]) / norm[4:]
    a1 = z @ W1.T + b1
    h1 = np.maximum(a1, 0.0)
    a2 = h1 @ W2.T + b2
    h2 = np.maximum(a2, 0.0)
    y = h2 @ W3 + b3
    N = len(t)
    d = y - np.asarray(t, np.float64)
    gy = 2.0 * d / N
    gW3 = h2.T @ gy
    gb3 = gy.sum()
    g2 = np.outer(gy, W3) * (a2 > 0)
    gW2 = g2.T @ h1
    gb2 = g2.sum(axis=0)
    g1 = (g2 @ W2) * (a1 > 0)
    gW1 = g1.T @ z
    gb1 = g1.sum(axis=0)
    return float(np.mean(d * d)), pack(gW1, gb1, gW2, gb2, gW3, gb3)


def normalization(x):
    x = np.asarray(x, np.float64)
    mu = x.mean(axis=0)
    var = (x * x).mean(axis=0) - mu * mu
    sd = np.where(var > 1e-12, np.sqrt(np.maximum(var, 0.0)), 1.0)
    return np.concatenate([mu, sd])
