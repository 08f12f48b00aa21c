"""CPU checks of the scheduler MLP's numpy reference (the checker of the HIP
kernels): its analytic gradient agrees with central finite differences, and a
zero-reward, zero-weight network has zero loss and gradient; and of the
training targets built from fixed-action runs."""
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


def test_targets_are_whole_evaluation_speedups():
    """Every recorded turn of a program, under whichever fixed action it was
    recorded, gets one row per action: log2 of the median evaluation time
    under action 0 over the median time under that action; at most
    `per_run` evenly spaced turns of a run are used."""
    import math

    import sched as S
    f = [[float(i)] * 32 for i in range(10)]
    runs = {0: [(100.0, f[:2]), (120.0, f[2:3]), (80.0, [])],
            1: [(50.0, f[3:4]), (50.0, [])],
            2: [(200.0, f[4:10])],
            3: [(100.0, [])]}
    rows, med = S.fixed_targets(runs, per_run=3)
    assert med == [100.0, 50.0, 200.0, 100.0]
    # turns used: f0 f1 (run 0), f2, f3, and 3 of f4..f9 evenly spaced: f4 f6 f8
    assert [r[0][0] for r in rows[::4]] == [0.0, 1.0, 2.0, 3.0, 4.0, 6.0, 8.0]
    assert [r[1] for r in rows[:4]] == [0, 1, 2, 3]
    assert [r[2] for r in rows[:4]] == [0.0, 1.0, -1.0, 0.0]
    assert all(r[2] == [0.0, 1.0, -1.0, 0.0][r[1]] for r in rows)
    assert math.isclose(sum(r[2] for r in rows), 0.0, abs_tol=1e-12)
