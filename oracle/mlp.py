"""CPU reference of the cooling-surrogate MLP (forward and backward).

TEST INFRASTRUCTURE ONLY: the checker for the HIP MLP kernels
(torch_amd/csrc/mlp.hip).  The network has no counterpart in the reference;
its arithmetic contract is defined here and pinned by finite differences
(``check_gradients``) and by torch autograd in tests/test_oracle_mlp.py.

Network: f (8 features, last two zero) -> 4 hidden layers of width 128 with
SiLU -> NOUT outputs.  Operands of every product are rounded to bfloat16
(round to nearest even) at the same points as the GPU kernels; accumulation
is float64 here (float32 on the GPU):

  forward   Z_L = bf16(H_{L-1}) . bf16(W_L)^T + b_L,   H_L = bf16(SiLU(Z_L))
            y   = bf16(H_4) . bf16(W_5)^T + b_5
  loss      0.5 / B * sum_b sum_o m_o (y_bo - t_bo)^2
  backward  dy = m (y - t) / B
            G_4 = bf16(dy) . bf16(W_5),  dW_5 = bf16(dy)^T . H_4,
            dZ_L = bf16(G_L * SiLU'(Z_L))
            G_{L-1} = dZ_L . bf16(W_L),  dW_L = dZ_L^T . bf16(H_{L-1}),
            db_L = sum_b dZ_L (dZ_L before rounding for the bias)
"""
from __future__ import annotations

import numpy as np

NIN, NH, NLAYER = 8, 128, 4


def bf16(x: np.ndarray) -> np.ndarray:
    """Round float64/float32 values to bfloat16 (nearest even), as float64."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    nan = np.isnan(f)
    rounding = ((u >> 16) & 1) + 0x7FFF
    u = ((u + rounding) >> 16) << 16
    out = u.astype(np.uint32).view(np.float32).astype(np.float64)
    out[nan] = np.nan
    return out


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def init_params(nout: int = 3, seed: int = 0) -> dict:
    rng = np.random.Generator(np.random.PCG64(seed))
    p = {}
    dims = [NIN] + [NH] * NLAYER + [nout]
    for L in range(1, NLAYER + 2):
        fan_in = dims[L - 1] if L > 1 else 6
        p[f"W{L}"] = rng.normal(0.0, np.sqrt(2.0 / fan_in), (dims[L], dims[L - 1]))
        if L == 1:
            p["W1"][:, 6:] = 0.0
        p[f"b{L}"] = np.zeros(dims[L])
    return p


def forward(p: dict, f: np.ndarray, round_bf16: bool = True):
    r = bf16 if round_bf16 else (lambda x: np.asarray(x, dtype=np.float64))
    h = r(f)
    cache = {"H0": h}
    for L in range(1, NLAYER + 1):
        z = h @ r(p[f"W{L}"]).T + p[f"b{L}"]
        cache[f"Z{L}"] = z
        h = r(silu(z))
        cache[f"H{L}"] = h
    y = h @ r(p[f"W{NLAYER + 1}"]).T + p[f"b{NLAYER + 1}"]
    return y, cache


def loss_and_grads(p: dict, f: np.ndarray, t: np.ndarray, m: np.ndarray, round_bf16: bool = True):
    r = bf16 if round_bf16 else (lambda x: np.asarray(x, dtype=np.float64))
    B = f.shape[0]
    y, c = forward(p, f, round_bf16)
    loss = 0.5 / B * np.sum(m * (y - t) ** 2)
    dy = m * (y - t) / B
    g = {}
    L5 = NLAYER + 1
    g[f"W{L5}"] = r(dy).T @ c[f"H{NLAYER}"]
    g[f"b{L5}"] = dy.sum(0)
    G = r(dy) @ r(p[f"W{L5}"])
    for L in range(NLAYER, 0, -1):
        dz_full = G * dsilu(c[f"Z{L}"])
        dz = r(dz_full)
        g[f"b{L}"] = dz_full.sum(0)
        g[f"W{L}"] = dz.T @ c[f"H{L - 1}"]
        if L > 1:
            G = dz @ r(p[f"W{L}"])
    return loss, g, y


def check_gradients(p: dict, f: np.ndarray, t: np.ndarray, m: np.ndarray, n_probe: int = 20, eps: float = 1e-6,
                    seed: int = 0) -> float:
    """Largest relative error of the analytic gradient (without bf16
    rounding) against central finite differences on random entries."""
    rng = np.random.Generator(np.random.PCG64(seed))
    _, g, _ = loss_and_grads(p, f, t, m, round_bf16=False)
    worst = 0.0
    for key in p:
        for _ in range(n_probe // 2 + 1):
            idx = tuple(rng.integers(0, s) for s in p[key].shape)
            if key == "W1" and idx[1] >= 6:
                continue
            old = p[key][idx]
            p[key][idx] = old + eps
            lp, _, _ = loss_and_grads(p, f, t, m, round_bf16=False)
            p[key][idx] = old - eps
            lm, _, _ = loss_and_grads(p, f, t, m, round_bf16=False)
            p[key][idx] = old
            fd = (lp - lm) / (2 * eps)
            an = g[key][idx]
            worst = max(worst, abs(fd - an) / max(abs(fd), abs(an), 1e-10))
    return worst
