"""MLP cost model: the fp64 CPU reference pinned by finite differences, and
the library's HIP forward/backward/Adam kernels checked against it."""
import numpy as np
import pytest

from oracle import mlp_oracle as M


def data(n, seed):
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(n, M.IN))
    Y = np.sin(X[:, 0]) + 0.5 * X[:, 1] * X[:, 2] - X[:, 3] ** 2 / 3
    return X, Y


def test_oracle_gradient_matches_finite_differences():
    X, Y = data(48, 1)
    p = M.init(2)
    g, _ = M.grad(p, X, Y)
    rng = np.random.default_rng(3)
    idx = np.concatenate([rng.choice(M.NPARAM, 60, replace=False), [M.NPARAM - 1, 0, 600, 4700]])
    h = 1e-6
    for i in idx:
        e = np.zeros(M.NPARAM)
        e[i] = h
        fd = (M.loss(p + e, X, Y) - M.loss(p - e, X, Y)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])


def test_oracle_adam_first_step_is_sign_step():
    # with bias correction the first Adam step moves every parameter with a
    # non-zero gradient by lr * sign(g) (up to eps)
    X, Y = data(32, 4)
    p = M.init(5)
    g, _ = M.grad(p, X, Y)
    p1, _ = M.adam(p, X, Y, 1, 1e-3)
    nz = np.abs(g) > 1e-3
    assert np.allclose((p - p1)[nz], 1e-3 * np.sign(g[nz]), rtol=1e-4, atol=1e-9)


@pytest.mark.gpu
def test_gpu_mlp_gradient_matches_oracle():
    import dumato_amd as dm
    X, Y = data(3000, 7)
    p = M.init(8)
    loss, g, out = dm.mlp_grad(p.astype(np.float32), X, Y)
    go, oo = M.grad(p.astype(np.float32).astype(np.float64), X.astype(np.float32), Y.astype(np.float32))
    assert np.allclose(out, oo, rtol=1e-4, atol=1e-4)
    assert abs(loss - M.loss(p.astype(np.float32), X.astype(np.float32), Y.astype(np.float32))) < 1e-4 * max(1, loss)
    # fp32 sums over 3000 samples: tolerance relative to the gradient scale
    scale = np.abs(go).max()
    assert np.max(np.abs(g - go)) <= 2e-4 * scale


@pytest.mark.gpu
def test_gpu_adam_matches_oracle():
    import dumato_amd as dm
    X, Y = data(1024, 9)
    p = M.init(10).astype(np.float32)
    steps = 25
    pg, lg = dm.mlp_adam(p, X, Y, steps, 1e-3)
    po, lo = M.adam(p.astype(np.float64), X.astype(np.float32), Y.astype(np.float32), steps, 1e-3)
    assert np.allclose(lg, lo, rtol=1e-3, atol=1e-5)
    assert np.max(np.abs(pg - po)) < 5e-4
    assert lg[-1] < lg[0]
