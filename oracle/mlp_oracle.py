"""CPU reference (float64, numpy) of the learned active-vertex filter: an 8 -> 64 -> 64 -> 1
perceptron with ReLU hidden layers and a logit output, trained with the weighted binary
cross-entropy L = mean[pw * y * softplus(-z) + (1 - y) * softplus(z)]. The filter has no
counterpart in the reference crate; this restatement is test infrastructure only (tests compare
the HIP forward and backward passes with it), and it is pinned by central finite differences
(tests/test_oracle.py). Parameter order: W1[8][64], b1[64], W2[64][64], b2[64], w3[64], b3."""
import numpy as np

NF, H = 8, 64
SIZES = [NF * H, H, H * H, H, H, 1]


def unpack(w):
    w = np.asarray(w, dtype=np.float64)
    out, o = [], 0
    for n in SIZES:
        out.append(w[o:o + n])
        o += n
    W1, b1, W2, b2, w3, b3 = out
    return W1.reshape(NF, H), b1, W2.reshape(H, H), b2, w3, b3[0]


def forward(w, X):
    W1, b1, W2, b2, w3, b3 = unpack(w)
    X = np.asarray(X, dtype=np.float64)
    a1 = X @ W1 + b1
    h1 = np.maximum(a1, 0.0)
    a2 = h1 @ W2 + b2
    h2 = np.maximum(a2, 0.0)
    z = h2 @ w3 + b3
    return z, (X, a1, h1, a2, h2)


def softplus(z):
    return np.where(z > 0, z + np.log1p(np.exp(-np.abs(z))), np.log1p(np.exp(-np.abs(z))))


def loss(w, X, y, pw=1.0):
    z, _ = forward(w, X)
    t = np.asarray(y, dtype=np.float64)
    return float(np.mean(pw * t * softplus(-z) + (1.0 - t) * softplus(z)))


def grad(w, X, y, pw=1.0):
    W1, b1, W2, b2, w3, b3 = unpack(w)
    z, (X, a1, h1, a2, h2) = forward(w, X)
    t = np.asarray(y, dtype=np.float64)
    sig = 1.0 / (1.0 + np.exp(-z))
    dz = (1.0 - t) * sig - pw * t * (1.0 - sig)
    n = len(z)
    d2 = np.where(a2 > 0, np.outer(dz, w3), 0.0)
    d1 = np.where(a1 > 0, d2 @ W2.T, 0.0)
    g = [X.T @ d1, d1.sum(0), h1.T @ d2, d2.sum(0), h2.T @ dz, np.array([dz.sum()])]
    return np.concatenate([x.reshape(-1) for x in g]) / n
