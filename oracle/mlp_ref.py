"""numpy restatement of the troubled-cell MLP and its input features.

TEST INFRASTRUCTURE ONLY: the checker of the HIP kernels in hopefoam_amd/csrc/mlp.hip and
limiter.hip (features, FP32 forward and gradients, BF16 inference).  Only tests/ load it.
The network and its features have no counterpart in the reference; the oracle is pinned by
finite differences (gradients), by closed-form cases (features) and by the IEEE bf16 rounding
rule (BF16 emulation).

The Adam step restates Kingma & Ba (ICLR 2015, Algorithm 1) in float32, as the training kernel
computes it; it is pinned by the closed forms of the bias-corrected update (the first step moves every
parameter with a non-zero gradient by lr in the direction of -sign(g); a constant gradient keeps doing so).

Parameter layout: W1 [64][nin], b1 [64], W2 [64][64], b2 [64], w3 [64], b3.
Loss: weighted binary cross-entropy with logits,
    L = mean_s [ pw y_s softplus(-z_s) + (1 - y_s) softplus(z_s) ].
"""
import numpy as np

H = 64


def unpack(p, nin):
    p = np.asarray(p, dtype=np.float64)
    o = 0
    W1 = p[o:o + H * nin].reshape(H, nin); o += H * nin
    b1 = p[o:o + H]; o += H
    W2 = p[o:o + H * H].reshape(H, H); o += H * H
    b2 = p[o:o + H]; o += H
    w3 = p[o:o + H]; o += H
    b3 = p[o]
    return W1, b1, W2, b2, w3, b3


def forward(p, X):
    X = np.asarray(X, dtype=np.float64)
    W1, b1, W2, b2, w3, b3 = unpack(p, X.shape[1])
    h1 = np.maximum(0.0, X @ W1.T + b1)
    h2 = np.maximum(0.0, h1 @ W2.T + b2)
    return h2 @ w3 + b3


def softplus(z):
    return np.where(z > 0, z + np.log1p(np.exp(-np.abs(z))), np.log1p(np.exp(-np.abs(z))))


def loss(p, X, y, pw=1.0):
    z = forward(p, X)
    return float(np.mean(pw * y * softplus(-z) + (1 - y) * softplus(z)))


def loss_grad(p, X, y, pw=1.0):
    """loss and its analytic gradient (float64)"""
    X = np.asarray(X, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    n, nin = X.shape
    W1, b1, W2, b2, w3, b3 = unpack(p, nin)
    a1 = X @ W1.T + b1
    h1 = np.maximum(0.0, a1)
    a2 = h1 @ W2.T + b2
    h2 = np.maximum(0.0, a2)
    z = h2 @ w3 + b3
    sig = 1.0 / (1.0 + np.exp(-z))
    L = float(np.mean(pw * y * softplus(-z) + (1 - y) * softplus(z)))
    dz = (pw * y * (sig - 1.0) + (1 - y) * sig) / n
    gw3 = h2.T @ dz
    gb3 = dz.sum()
    d2 = np.outer(dz, w3) * (a2 > 0)
    gW2 = d2.T @ h1
    gb2 = d2.sum(0)
    d1 = (d2 @ W2) * (a1 > 0)
    gW1 = d1.T @ X
    gb1 = d1.sum(0)
    return L, np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gw3, [gb3]])


def bf16(x):
    """round float32 values to bfloat16 (nearest, ties to even), returned as float32"""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def forward_bf16(p, X):
    """BF16 inference: inputs, W1, W2 and the layer-1 activations in bf16, sums in float64"""
    X = np.asarray(X, dtype=np.float32)
    W1, b1, W2, b2, w3, b3 = unpack(p, X.shape[1])
    xb = bf16(X).astype(np.float64)
    h1 = np.maximum(0.0, xb @ bf16(W1.astype(np.float32)).astype(np.float64).T + b1.astype(np.float32))
    h1b = bf16(h1.astype(np.float32)).astype(np.float64)
    h2 = np.maximum(0.0, h1b @ bf16(W2.astype(np.float32)).astype(np.float64).T + b2.astype(np.float32))
    return h2 @ w3.astype(np.float32) + np.float32(b3)


def features(N, q, mean, nb_mean, invV, mid, h, M):
    """inputs of every cell: q [4][K][Np]; mean [4][K] own means; nb_mean [4][K][3] neighbour (or ghost)
    means; invV [Np][Np]; mid [4][K][3] P1-part values at the edge midpoints; h [K] longest edges; M the
    TVB constant.  Returns X [K][2 Np + 8] (float32)."""
    K = q.shape[1]
    out = []
    for v in (0, 3):
        c = q[v] @ invV.T
        out.append(c / (np.abs(c[:, :1]) + 1e-12))
    for v in (0, 3):
        u0 = mean[v][:, None]
        out.append((nb_mean[v] - u0) / (np.abs(u0) + 1e-12))
    for v in (0, 3):
        dmax = np.max(np.abs(mid[v] - mean[v][:, None]), axis=1)
        out.append(np.clip(np.log10(dmax / (M * h * h) + 1e-30), -10.0, 10.0)[:, None])
    return np.concatenate(out, axis=1).astype(np.float32).reshape(K, -1)


def adam_step(p, m, v, g, t, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam step (Kingma & Ba 2015, Algorithm 1) in float32: m = b1 m + (1 - b1) g,
    v = b2 v + (1 - b2) g^2, p -= lr mhat / (sqrt(vhat) + eps) with mhat = m / (1 - b1^t),
    vhat = v / (1 - b2^t); t counts steps from 1.  Returns new (p, m, v)."""
    f = np.float32
    p, m, v, g = (np.asarray(a, dtype=f) for a in (p, m, v, g))
    b1, b2 = f(beta1), f(beta2)
    m = b1 * m + (f(1) - b1) * g
    v = b2 * v + (f(1) - b2) * g * g
    mh = m / (f(1) - b1 ** f(t))
    vh = v / (f(1) - b2 ** f(t))
    return p - f(lr) * mh / (np.sqrt(vh) + f(eps)), m, v
