"""The trained cooling surrogate (torch_amd/data/surrogate_cooling.npz,
written by `python -m torch_amd.train_surrogate` on the GPU) against the FP64
subcycled integrator on fresh cells of the cooling workload (seed 11 family,
disjoint from the training seeds 1000+).

What the network gains over the method without it: the surrogate step
(FP64 prepass, BF16-MFMA MLP, FP64 fallback behind the 1% error gate) must
be faster than the subcycled FP64 integrator on the same cells while its
temperatures stay within 1% median and 5% p99 of the FP64 ones."""
import os

import pytest

pytestmark = pytest.mark.gpu
WEIGHTS = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "torch_amd", "data",
                       "surrogate_cooling.npz")


@pytest.mark.skipif(not os.path.exists(WEIGHTS), reason="surrogate weights not trained yet")
def test_surrogate_accuracy_and_gain(gpu):
    import torch

    from torch_amd import surrogate as sg
    from torch_amd.train_surrogate import workload_inputs

    cool = gpu.Cooling(gpu.thermo_params())
    mlp = sg.attach(cool, WEIGHTS, gate=0.01)
    warm = workload_inputs(1 << 10, seed=12)
    cool.thermo_cells(warm)
    sg.thermo_cells_surrogate(cool, warm)
    torch.cuda.synchronize()
    inp = workload_inputs(1 << 16, seed=13)
    r = sg.compare(cool, inp)
    print(r)
    assert r["rel_T_err_all"]["median"] <= 0.01, r
    assert r["rel_T_err_all"]["p99"] <= 0.05, r
    assert r["mlp_accepted"] > 0
    assert r["surrogate_ms"] < r["fp64_ms"], r
    del mlp
