"""The trained cooling surrogate (torch_amd/data/surrogate_cooling.npz,
written by `python -m torch_amd.train_surrogate` on the GPU) against the FP64
subcycled integrator on fresh cells of the cooling workload (seed 11 family,
disjoint from the training seeds 1000+).

What the network gains over the methods without it:
* against the subcycled FP64 integrator: the surrogate step (FP64 prepass
  with the coarse end-temperature estimate, BF16-MFMA MLP, FP64 fallback
  behind the 1% error gate) must be at least 10x faster on the same cells
  while its temperatures stay within 1% median and 5% p99 of the FP64 ones;
* against the coarse estimate alone (the same step with a zero network,
  T = Tc): the trained network must cut the p99 temperature error of the
  cells it handles at least five-fold."""
import os

import numpy as np
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
    r, ref = sg.compare(cool, inp, return_fp64=True)
    print(r)
    assert r["rel_T_err_all"]["median"] <= 0.01, r
    assert r["rel_T_err_all"]["p99"] <= 0.05, r
    assert r["rel_T_err_mlp_cells"]["median"] <= 0.01 and r["rel_T_err_mlp_cells"]["p99"] <= 0.05, r
    assert r["mlp_accepted"] > 0.5 * r["mlp_queued"], r
    assert 10.0 * r["surrogate_ms"] < r["fp64_ms"], r

    # the coarse estimate alone: a zero network gives T = Tc (and an
    # estimated error of sigma < 1%, so every queued cell is accepted)
    d = sg.load(WEIGHTS)
    zero = sg.MLP(max_batch=1024)
    zero.set_params(np.zeros(sg.NW, np.float32), np.zeros(sg.NB, np.float32))
    sg.set_surrogate(cool, zero, sg.make_cfg(list(d["mu"]), list(d["isd"]), d["lo"], d["hi"], gate=0.01,
                                             nexact=int(d["nexact"])))
    out, path, _ = sg.thermo_cells_surrogate(cool, inp)
    rel = ((out[4] / ref[4]) - 1.0).abs()[path == 1]
    coarse_p99 = float(torch.quantile(rel, 0.99))
    print("coarse estimate alone: p99", coarse_p99, "median", float(rel.median()))
    assert r["rel_T_err_mlp_cells"]["p99"] * 5.0 < coarse_p99, (r, coarse_p99)
    del mlp, zero
