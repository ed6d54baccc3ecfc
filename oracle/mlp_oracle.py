"""ORACLE -- TEST INFRASTRUCTURE ONLY.

float64 numpy restatement of the surrogate MLP the GPU trains
(seisfem_amd/csrc/sfg_mlp.hip): Linear + GELU (tanh form) hidden layers, a
Linear output, mean-squared-error loss, Adam.  Where the GPU rounds to bf16
(GEMM operands: activations, weights, output-layer gradients) the oracle
rounds the same values to bf16 (round to nearest even) and then computes in
float64, so GPU and oracle differ only by FP32 accumulation order.

Pinned by central finite differences of the loss (``check_gradients`` with
bf16 rounding switched off), which fixes every term of the analytic backward
pass.  The reference has no learned component; nothing here restates it.
"""
import numpy as np

K_GELU = np.sqrt(2.0 / np.pi)


def bf16(x):
    """Round float values to bf16 (nearest even), returned as float64."""
    a = np.ascontiguousarray(x, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def gelu(x):
    return 0.5 * x * (1.0 + np.tanh(K_GELU * (x + 0.044715 * x ** 3)))


def gelu_grad(x):
    u = K_GELU * (x + 0.044715 * x ** 3)
    th = np.tanh(u)
    return 0.5 * (1.0 + th) + 0.5 * x * (1.0 - th * th) * K_GELU * (1.0 + 3 * 0.044715 * x * x)


def forward(Ws, bs, X, round_bf16=True):
    """Returns (y, cache).  Ws[l]: [out][in], bs[l]: [out]."""
    r = bf16 if round_bf16 else (lambda a: np.asarray(a, np.float64))
    h = r(X)
    hs, zs = [h], []
    for l, (W, b) in enumerate(zip(Ws, bs)):
        z = h @ r(W).T + b
        zs.append(z)
        if l < len(Ws) - 1:
            h = r(gelu(z))
            hs.append(h)
    return zs[-1], (hs, zs)


def loss_and_grads(Ws, bs, X, T, round_bf16=True):
    r = bf16 if round_bf16 else (lambda a: np.asarray(a, np.float64))
    y, (hs, zs) = forward(Ws, bs, X, round_bf16)
    B, D = y.shape
    loss = float(np.mean((y - T) ** 2))
    dz = 2.0 * (y - T) / (B * D)
    gW, gb = [None] * len(Ws), [None] * len(Ws)
    for l in range(len(Ws) - 1, -1, -1):
        dzr = r(dz)
        gW[l] = dzr.T @ hs[l]
        gb[l] = dz.sum(axis=0)
        if l > 0:
            dh = dzr @ r(Ws[l])
            dz = dh * gelu_grad(zs[l - 1])
    return loss, gW, gb


def adam_step(w, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    w = w - lr * (m / (1 - b1 ** step)) / (np.sqrt(v / (1 - b2 ** step)) + eps)
    return w, m, v


def check_gradients(Ws, bs, X, T, eps=1e-6, samples=20, seed=0):
    """Max relative error of the analytic gradient against central
    differences of the float64 loss, over random weight entries."""
    rng = np.random.default_rng(seed)
    _, gW, gb = loss_and_grads(Ws, bs, X, T, round_bf16=False)
    worst = 0.0
    for l in range(len(Ws)):
        for _ in range(samples):
            i, j = rng.integers(0, Ws[l].shape[0]), rng.integers(0, Ws[l].shape[1])
            Wp = [w.copy() for w in Ws]
            Wm = [w.copy() for w in Ws]
            Wp[l][i, j] += eps
            Wm[l][i, j] -= eps
            fd = (loss_and_grads(Wp, bs, X, T, False)[0] - loss_and_grads(Wm, bs, X, T, False)[0]) / (2 * eps)
            worst = max(worst, abs(fd - gW[l][i, j]) / max(abs(fd), 1e-12))
            k = rng.integers(0, bs[l].shape[0])
            bp = [b.copy() for b in bs]
            bm = [b.copy() for b in bs]
            bp[l][k] += eps
            bm[l][k] -= eps
            fd = (loss_and_grads(Ws, bp, X, T, False)[0] - loss_and_grads(Ws, bm, X, T, False)[0]) / (2 * eps)
            worst = max(worst, abs(fd - gb[l][k]) / max(abs(fd), 1e-12))
    return worst
