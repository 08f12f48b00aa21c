"""float32 numpy reference of the scheduler MLP (32 -> 64 -> 64 -> 4, ReLU),
used to check the HIP forward/backward kernels.  Parameters are packed
W1[32][64] b1[64] W2[64][64] b2[64] W3[64][4] b3[4]; loss per sample is
0.5 * (y[action] - reward)^2, averaged over the batch."""
import numpy as np

F, H, K = 32, 64, 4
SIZES = [F * H, H, H * H, H, H * K, K]


def unpack(p):
    p = np.asarray(p, dtype=np.float32)
    out, o = [], 0
    for s in SIZES:
        out.append(p[o:o + s])
        o += s
    W1, b1, W2, b2, W3, b3 = out
    return W1.reshape(F, H), b1, W2.reshape(H, H), b2, W3.reshape(H, K), b3


def forward(p, X):
    W1, b1, W2, b2, W3, b3 = unpack(p)
    X = np.asarray(X, dtype=np.float32)
    Z1 = X @ W1 + b1
    H1 = np.maximum(Z1, 0)
    Z2 = H1 @ W2 + b2
    H2 = np.maximum(Z2, 0)
    Y = H2 @ W3 + b3
    return Y, (X, H1, H2)


def loss_and_grad(p, X, A, R):
    W1, b1, W2, b2, W3, b3 = unpack(p)
    Y, (X, H1, H2) = forward(p, X)
    n = X.shape[0]
    A = np.asarray(A)
    R = np.asarray(R, dtype=np.float32)
    dY = np.zeros_like(Y)
    dY[np.arange(n), A] = Y[np.arange(n), A] - R
    loss = float(0.5 * (dY ** 2).sum() / n)
    dH2 = (dY @ W3.T) * (H2 > 0)
    dH1 = (dH2 @ W2.T) * (H1 > 0)
    gW3 = H2.T @ dY / n
    gb3 = dY.sum(0) / n
    gW2 = H1.T @ dH2 / n
    gb2 = dH2.sum(0) / n
    gW1 = X.T @ dH1 / n
    gb1 = dH1.sum(0) / n
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])
    return loss, g


def to_half_bits(p):
    return np.asarray(p, dtype=np.float16).view(np.uint16)


def from_half_bits(bits):
    return np.asarray(bits, dtype=np.uint16).view(np.float16).astype(np.float32)
