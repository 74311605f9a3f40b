"""FP64 reference of the learned-guess MLP and its training step.

TEST INFRASTRUCTURE ONLY: the checker of the HIP kernels in
plugin_navierstokes_amd/csrc/nsfv1_mlp.hip (forward, backward, Adam,
features, per-step scale, guess and target).  Never imported by the
product path.

The network has no counterpart in the reference plugin.  Its arithmetic:
    x  = bf16(features)                    [n, 32]  (24 used, 8 zero)
    z1 = W1 x + b1,  h1 = bf16(relu(z1))   [n, 128]
    z2 = W2 h1 + b2, h2 = bf16(relu(z2))   [n, 128]
    y  = W3 h2 + b3                        [n, 4]
with W* = bf16(master FP32 weights) and biases in FP32; every sum is
exact here (FP64), so the GPU differs only by its FP32 accumulation (and
the bf16 rounding it may then take differently).  Loss = mean over n x 4
of (y - t)^2.  Gradients treat the bf16 roundings as identity
(straight-through) and use the rounded weights, as the GPU does.
"""
from __future__ import annotations

import numpy as np

IN, KIN, H, OUT, KOUT = 24, 32, 128, 4, 16
CLAMP = 64.0
OFF_W1 = 0
OFF_B1 = OFF_W1 + H * KIN
OFF_W2 = OFF_B1 + H
OFF_B2 = OFF_W2 + H * H
OFF_W3 = OFF_B2 + H
OFF_B3 = OFF_W3 + KOUT * H
NPARAM = OFF_B3 + KOUT


def bf16(a) -> np.ndarray:
    """Round FP32 values to bfloat16 (round to nearest even), returned as FP64."""
    a32 = np.asarray(a, dtype=np.float32)
    u = a32.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def unpack(P: np.ndarray):
    P = np.asarray(P, dtype=np.float64)
    W1 = P[OFF_W1:OFF_B1].reshape(H, KIN)
    b1 = P[OFF_B1:OFF_W2]
    W2 = P[OFF_W2:OFF_B2].reshape(H, H)
    b2 = P[OFF_B2:OFF_W3]
    W3 = P[OFF_W3:OFF_B3].reshape(KOUT, H)[:OUT]
    b3 = P[OFF_B3:OFF_B3 + OUT]
    return W1, b1, W2, b2, W3, b3


def init_params(seed: int = 0) -> np.ndarray:
    """He-uniform initialisation with zero padding (FP32 values)."""
    rng = np.random.default_rng(seed)
    P = np.zeros(NPARAM, dtype=np.float32)
    W1 = np.zeros((H, KIN), np.float32)
    W1[:, :IN] = rng.uniform(-1, 1, (H, IN)) * np.sqrt(6.0 / IN)
    W2 = rng.uniform(-1, 1, (H, H)) * np.sqrt(6.0 / H)
    W3 = np.zeros((KOUT, H), np.float32)
    W3[:OUT] = rng.uniform(-1, 1, (OUT, H)) * np.sqrt(6.0 / H) * 0.1
    P[OFF_W1:OFF_B1] = W1.ravel()
    P[OFF_W2:OFF_B2] = W2.ravel()
    P[OFF_W3:OFF_B3] = W3.ravel()
    return P


def forward(P, X, rounding: bool = True):
    """Returns (y, cache).  rounding=False gives the plain FP64 network (used
    to check the gradients by finite differences)."""
    r = bf16 if rounding else (lambda a: np.asarray(a, dtype=np.float64))
    W1, b1, W2, b2, W3, b3 = unpack(P)
    W1, W2, W3 = r(W1), r(W2), r(W3)
    x = r(np.asarray(X, dtype=np.float64)[:, :KIN])
    z1 = x @ W1.T + b1
    h1 = r(np.maximum(z1, 0.0))
    z2 = h1 @ W2.T + b2
    h2 = r(np.maximum(z2, 0.0))
    y = h2 @ W3.T + b3
    return y, dict(x=x, z1=z1, h1=h1, z2=z2, h2=h2, W1=W1, W2=W2, W3=W3)


def output_scale(P, X):
    """Sum of |terms| of each output (the magnitude scale of its rounding)."""
    W1, b1, W2, b2, W3, b3 = unpack(P)
    _, c = forward(P, X)
    return np.abs(c["h2"]) @ np.abs(c["W3"]).T + np.abs(b3)


def loss_and_grad(P, X, T, rounding: bool = True):
    y, c = forward(P, X, rounding)
    n = y.shape[0]
    e = y - T
    loss = float(np.mean(e * e))
    dy = 2.0 * e / (n * OUT)
    dz2 = (dy @ c["W3"]) * (c["z2"] > 0)
    dz1 = (dz2 @ c["W2"]) * (c["z1"] > 0)
    g = np.zeros(NPARAM)
    g[OFF_W1:OFF_B1] = (dz1.T @ c["x"]).ravel()
    g[OFF_B1:OFF_W2] = dz1.sum(0)
    g[OFF_W2:OFF_B2] = (dz2.T @ c["h1"]).ravel()
    g[OFF_B2:OFF_W3] = dz2.sum(0)
    gw3 = np.zeros((KOUT, H))
    gw3[:OUT] = dy.T @ c["h2"]
    g[OFF_W3:OFF_B3] = gw3.ravel()
    g[OFF_B3:OFF_B3 + OUT] = dy.sum(0)
    # padding entries carry no gradient
    g[OFF_W1:OFF_B1].reshape(H, KIN)[:, IN:] = 0.0
    return loss, g, dy


def adam(P, M, V, G, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One bias-corrected Adam step; padding stays zero.  Returns new (P, M, V)."""
    pad = np.zeros(NPARAM, dtype=bool)
    pad[OFF_W1:OFF_B1] = (np.arange(H * KIN) % KIN) >= IN
    pad[OFF_W3:OFF_B3] = (np.arange(KOUT * H) // H) >= OUT
    pad[OFF_B3:] = np.arange(KOUT) >= OUT
    M = b1 * M + (1 - b1) * G
    V = b2 * V + (1 - b2) * G * G
    c1, c2 = 1.0 / (1 - b1 ** step), 1.0 / (1 - b2 ** step)
    P = P - lr * (M * c1) / (np.sqrt(V * c2) + eps)
    P[pad] = 0.0
    return P, M, V


def differences(u, u1, u2, u3):
    """Backward differences d1, d2, d3 of the last four states, as nested
    differences: d1 = u - u1, e1 = u1 - u2, d2 = d1 - e1, d3 = d2 - (e1 - (u2 - u3))."""
    d1, e1 = u - u1, u1 - u2
    d2 = d1 - e1
    return d1, d2, d2 - (e1 - (u2 - u3))


def step_scale(u, u1, u2, u3) -> float:
    """sigma = max(rms of d3 over the velocity components of all nodes, 1e-30)."""
    d3 = differences(u, u1, u2, u3)[2][:, :3]
    r = float(np.sqrt(np.sum(d3 * d3) / (3.0 * u.shape[0]))) if u.shape[0] else 0.0
    return max(r, 1e-30)


def _f32_scaled(a, s):
    """clamp(float32(a) * float32(s), +-CLAMP) in FP32 arithmetic."""
    p = np.asarray(a, dtype=np.float64).astype(np.float32) * np.asarray(s, dtype=np.float32)
    return np.clip(p, np.float32(-CLAMP), np.float32(CLAMP)).astype(np.float64)


def features(rowptr, colidx, u, u1, u2, u3, scale, sigma=None):
    """Per-node features
        (u, d1/sigma, d2/sigma, d3/sigma, ring(u) - u, (ring(d1) - d1)/sigma) * scale,
    each clamped to +-CLAMP, ring() the mean over the node's 1-ring taken
    from the CSR pattern's column blocks of component 0.  Returns (X, sigma)."""
    if sigma is None:
        sigma = step_scale(u, u1, u2, u3)
    d1, d2, d3 = differences(u, u1, u2, u3)
    isg = 1.0 / sigma
    scale = np.asarray(scale, dtype=np.float32)
    nv = u.shape[0]
    X = np.zeros((nv, KIN))
    for v in range(nv):
        cols = colidx[rowptr[4 * v]:rowptr[4 * v + 1]:4] // 4
        ring = [w for w in cols if w != v]
        m0 = u[ring].mean(0) - u[v] if ring else np.zeros(4)
        m1 = (d1[ring].mean(0) - d1[v]) * isg if ring else np.zeros(4)
        raw = [u[v], d1[v] * isg, d2[v] * isg, d3[v] * isg, m0, m1]
        for g, a in enumerate(raw):
            X[v, 4 * g:4 * g + 4] = _f32_scaled(a, scale[4 * g:4 * g + 4])
    return X, sigma


def quadratic(u, u1, u2):
    """Quadratic extrapolation u + d1 + d2 (= 3 u - 3 u1 + u2, the base of the guess)."""
    d1 = u - u1
    return u + d1 + (d1 - (u1 - u2))


def guess(u, u1, u2, y, out_scale, sigma):
    return quadratic(u, u1, u2) + np.asarray(y, dtype=np.float64) * np.asarray(out_scale, dtype=np.float32) * sigma


def target(u_next, u, u1, u2, out_scale, sigma):
    """clamp(((u^{n+1} - u) - d1 - d2) / (sigma * out_scale)) as FP32; 0 where out_scale is 0."""
    os = np.asarray(out_scale, dtype=np.float32).astype(np.float64)
    d1 = u - u1
    r = ((u_next - u) - d1) - (d1 - (u1 - u2))
    with np.errstate(divide="ignore", invalid="ignore"):
        t = np.where(os > 0, r / (sigma * np.where(os > 0, os, 1.0)), 0.0)
    return np.clip(t.astype(np.float32), np.float32(-CLAMP), np.float32(CLAMP)).astype(np.float64)
