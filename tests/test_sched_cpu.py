"""CPU checks of the scheduler MLP's numpy reference (the checker of the HIP
kernels): its analytic gradient agrees with central finite differences, and a
zero-reward, zero-weight network has zero loss and gradient."""
import numpy as np

import mlp_ref


def test_reference_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    n = 16
    p = rng.normal(0, 0.3, sum(mlp_ref.SIZES)).astype(np.float64)
    X = rng.random((n, mlp_ref.F))
    A = rng.integers(0, mlp_ref.K, n)
    R = rng.normal(0, 1, n)

    def loss64(q):
        W1, b1, W2, b2, W3, b3 = [a.astype(np.float64) for a in mlp_ref.unpack(q)]
        H1 = np.maximum(X @ W1 + b1, 0)
        H2 = np.maximum(H1 @ W2 + b2, 0)
        Y = H2 @ W3 + b3
        d = Y[np.arange(n), A] - R
        return 0.5 * float((d ** 2).sum()) / n

    _, g = mlp_ref.loss_and_grad(p, X, A, R)
    idx = rng.choice(len(p), 60, replace=False)
    eps = 1e-3
    for i in idx:
        q = p.copy(); q[i] += eps
        lp = loss64(q)
        q[i] -= 2 * eps
        lm = loss64(q)
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - g[i]) <= 2e-3 * max(1.0, abs(fd)), (i, fd, g[i])


def test_zero_network():
    p = np.zeros(sum(mlp_ref.SIZES), dtype=np.float32)
    X = np.ones((4, mlp_ref.F), dtype=np.float32)
    loss, g = mlp_ref.loss_and_grad(p, X, [0, 1, 2, 3], [0, 0, 0, 0])
    assert loss == 0.0 and not g.any()
