"""CPU reference (float64 numpy) of the cost-model MLP: forward, loss
gradient and Adam.  TEST INFRASTRUCTURE ONLY: the checker for the library's
HIP MLP kernels (dumato_amd/csrc/mlp.hip).  No reference-repository
counterpart exists (the learned scheduler is new); it is pinned by central
finite differences of the loss (tests/test_mlp_oracle.py) and by Adam's
published update rule (Kingma & Ba 2015, Algorithm 1).

Parameter layout (4801 values): W1[8][64], b1[64], W2[64][64], b2[64],
W3[64], b3; hidden layers use ReLU; loss = mean_b (f(x_b) - y_b)^2.
"""
import numpy as np

IN, H = 8, 64
NPARAM = IN * H + H + H * H + H + H + 1


def unpack(p):
    p = np.asarray(p, np.float64)
    o = 0
    W1 = p[o:o + IN * H].reshape(IN, H); o += IN * H
    b1 = p[o:o + H]; o += H
    W2 = p[o:o + H * H].reshape(H, H); o += H * H
    b2 = p[o:o + H]; o += H
    W3 = p[o:o + H]; o += H
    b3 = p[o]
    return W1, b1, W2, b2, W3, b3


def forward(p, X):
    W1, b1, W2, b2, W3, b3 = unpack(p)
    X = np.asarray(X, np.float64)
    z1 = X @ W1 + b1
    h1 = np.maximum(z1, 0)
    z2 = h1 @ W2 + b2
    h2 = np.maximum(z2, 0)
    return h2 @ W3 + b3, (z1, h1, z2, h2)


def loss(p, X, Y):
    out, _ = forward(p, X)
    return float(np.mean((out - np.asarray(Y, np.float64)) ** 2))


def grad(p, X, Y):
    W1, b1, W2, b2, W3, b3 = unpack(p)
    X = np.asarray(X, np.float64)
    out, (z1, h1, z2, h2) = forward(p, X)
    n = len(out)
    g = 2.0 * (out - np.asarray(Y, np.float64)) / n
    dW3 = h2.T @ g
    db3 = g.sum()
    dz2 = np.outer(g, W3) * (z2 > 0)
    dW2 = h1.T @ dz2
    db2 = dz2.sum(0)
    dz1 = (dz2 @ W2.T) * (z1 > 0)
    dW1 = X.T @ dz1
    db1 = dz1.sum(0)
    return np.concatenate([dW1.ravel(), db1, dW2.ravel(), db2, dW3, [db3]]), out


def adam(p, X, Y, steps, lr, b1=0.9, b2=0.999, eps=1e-8):
    p = np.array(p, np.float64)
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    losses = []
    for t in range(1, steps + 1):
        losses.append(loss(p, X, Y))
        g, _ = grad(p, X, Y)
        m = b1 * m + (1 - b1) * g
        v = b2 * v + (1 - b2) * g * g
        mh = m / (1 - b1 ** t)
        vh = v / (1 - b2 ** t)
        p = p - lr * mh / (np.sqrt(vh) + eps)
    return p, np.array(losses)


def init(seed):
    rng = np.random.default_rng(seed)
    W1 = rng.normal(0, np.sqrt(2 / IN), (IN, H))
    W2 = rng.normal(0, np.sqrt(2 / H), (H, H))
    W3 = rng.normal(0, np.sqrt(1 / H), H)
    return np.concatenate([W1.ravel(), rng.normal(0, 0.1, H), W2.ravel(), rng.normal(0, 0.1, H), W3, [0.05]])
