"""BF16-MFMA MLP kernels and the surrogate cooling path on the GPU.

Gradients of the library's forward/backward kernels are compared with the
CPU reference oracle/mlp.py, which rounds the same operands to bfloat16 and
accumulates in float64 (the GPU accumulates in float32): tolerance 1e-3
relative L2 per parameter tensor.  The fused inference kernel is compared
with the same reference forward (1e-3 relative)."""
import numpy as np
import pytest

from oracle import mlp as omlp
from torch_amd import cooling_inputs

pytestmark = pytest.mark.gpu


def _setup(seed=0, B=512):
    import torch

    from torch_amd import surrogate as sg

    W, b = sg.init_params(seed)
    rng = np.random.Generator(np.random.PCG64(seed + 1))
    b = rng.normal(0, 0.1, b.shape).astype(np.float32)  # non-zero biases exercise the bias paths
    f = np.zeros((B, 8), np.float32)
    f[:, :6] = rng.normal(0, 1, (B, 6))
    t = rng.normal(0, 1, (B, 3)).astype(np.float32)
    m = sg.MLP(max_batch=4096)
    m.set_params(W, b)
    Ws, bs = sg.unpack(W, b)
    p = {f"W{L + 1}": Ws[L] for L in range(5)}
    p.update({f"b{L + 1}": bs[L] for L in range(5)})
    return m, p, f, t, torch.from_numpy(f).cuda(), torch.from_numpy(t).cuda()


def test_mlp_gradients_match_cpu_reference(gpu):
    from torch_amd import surrogate as sg

    m, p, f, t, fd, td = _setup()
    loss, gW, gb = m.loss_grad(fd, td, weights=(1.0, 0.5, 0.25))
    wts = np.array([1.0, 0.5, 0.25])
    ref_loss, g, _ = omlp.loss_and_grads(p, f.astype(np.float64), t.astype(np.float64), wts)
    assert abs(loss - ref_loss) / ref_loss < 1e-4
    Ws, bs = sg.unpack(gW, gb)
    for L in range(5):
        for name, got in ((f"W{L + 1}", Ws[L]), (f"b{L + 1}", bs[L])):
            ref = g[name]
            err = np.linalg.norm(got - ref) / max(np.linalg.norm(ref), 1e-30)
            assert err < 1e-3, (name, err)


def test_mlp_fused_inference_matches_reference(gpu):
    m, p, f, t, fd, td = _setup(seed=3, B=1000)  # ragged: not a multiple of 32
    y = m.forward(fd).cpu().numpy()
    yr, _ = omlp.forward(p, f.astype(np.float64))
    err = np.abs(y - yr).max() / np.abs(yr).max()
    assert err < 1e-3, err


def test_mlp_training_reduces_loss_and_is_deterministic(gpu):
    import torch

    from torch_amd import surrogate as sg

    rng = np.random.Generator(np.random.PCG64(5))
    B = 4096
    f = np.zeros((B, 8), np.float32)
    f[:, :6] = rng.uniform(-1, 1, (B, 6))
    tgt = np.zeros((B, 3), np.float32)
    tgt[:, 0] = np.sin(2 * f[:, 0]) * f[:, 1] + 0.3 * f[:, 2]
    fd, td = torch.from_numpy(f).cuda(), torch.from_numpy(tgt).cuda()
    finals = []
    for _ in range(2):
        m = sg.MLP(max_batch=B)
        m.set_params(*sg.init_params(7))
        first = m.train_step(fd, td, 1e-3, err_head=False, want_loss=True)
        for _ in range(300):
            last = m.train_step(fd, td, 1e-3, err_head=False, want_loss=True)
        finals.append((first, last, m.get_params()))
        m.close()
    assert finals[0][1] < 0.05 * finals[0][0], (finals[0][0], finals[0][1])
    np.testing.assert_array_equal(finals[0][2][0], finals[1][2][0])  # bitwise reproducible training


def _cooling_inputs(orc, n, seed):
    c = orc.converter(1e-20, 1e-28, 1e11)
    tc = lambda v, m, l, t: orc.to_code(c, v, m, l, t)  # noqa: E731
    return cooling_inputs.to_code_inputs(cooling_inputs.sample_cgs(n, seed=seed), tc)


def test_surrogate_path_all_fallback_equals_fp64(gpu, orc):
    """With the gate closed every queued cell takes the FP64 fallback: the
    surrogate step must then reproduce the FP64 integrator bitwise, and the
    exactly-integrated cells (<= nexact subcycles) always do."""
    import torch

    from torch_amd import surrogate as sg

    inp = torch.from_numpy(np.ascontiguousarray(_cooling_inputs(orc, 20000, 17).T)).cuda()
    cool = gpu.Cooling(gpu.thermo_params())
    ref, _ = cool.thermo_cells(inp)
    m = sg.MLP(max_batch=1024)
    m.set_params(*sg.init_params(1))
    big = [1e30] * 6
    sg.set_surrogate(cool, m, sg.make_cfg([0] * 8, [1] * 8, [-x for x in big], big, gate=-1.0, nexact=4))
    out, path, (ns, nf) = sg.thermo_cells_surrogate(cool, inp)
    torch.testing.assert_close(out, ref, rtol=0, atol=0)
    p = path.cpu().numpy()
    assert ns == nf == int((p == 2).sum()) and ns > 0
    nsub = ref[3].cpu().numpy()
    assert (p[nsub < 4] == 0).all() and (p[nsub >= 4] == 2).all()
