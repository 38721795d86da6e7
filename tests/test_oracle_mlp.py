"""Pins the CPU MLP reference (oracle/mlp.py) itself: its analytic gradient
against central finite differences (no bf16 rounding) and against torch
autograd in float64, and its bf16 rounding against torch's bfloat16."""
import numpy as np
import torch

from oracle import mlp as omlp


def _data(B=64, seed=0):
    rng = np.random.Generator(np.random.PCG64(seed))
    f = np.zeros((B, 8))
    f[:, :6] = rng.normal(0, 1, (B, 6))
    t = rng.normal(0, 1, (B, 3))
    return f, t


def test_gradients_vs_finite_differences():
    p = omlp.init_params(seed=1)
    for k in p:
        if k.startswith("b"):
            p[k] = np.random.Generator(np.random.PCG64(2)).normal(0, 0.1, p[k].shape)
    f, t = _data()
    assert omlp.check_gradients(p, f, t, np.array([1.0, 0.5, 0.25]), n_probe=30, eps=1e-4) < 1e-5


def test_gradients_vs_torch_autograd():
    p = omlp.init_params(seed=4)
    f, t = _data(seed=5)
    m = np.array([1.0, 0.3, 2.0])
    loss, g, _ = omlp.loss_and_grads(p, f, t, m, round_bf16=False)
    tp = {k: torch.tensor(v, requires_grad=True) for k, v in p.items()}
    h = torch.tensor(f)
    for L in range(1, 5):
        h = torch.nn.functional.silu(h @ tp[f"W{L}"].T + tp[f"b{L}"])
    y = h @ tp["W5"].T + tp["b5"]
    tl = 0.5 / f.shape[0] * (torch.tensor(m) * (y - torch.tensor(t)) ** 2).sum()
    tl.backward()
    assert abs(tl.item() - loss) < 1e-12 * max(1.0, abs(loss))
    for k in p:
        np.testing.assert_allclose(g[k], tp[k].grad.numpy(), rtol=1e-10, atol=1e-14)


def test_bf16_rounding_matches_torch():
    x = np.random.Generator(np.random.PCG64(9)).normal(0, 10, 100000)
    ours = omlp.bf16(x)
    ref = torch.tensor(x, dtype=torch.float32).to(torch.bfloat16).to(torch.float64).numpy()
    np.testing.assert_array_equal(ours, ref)
