"""The cloud-trained cooling surrogate (torch_amd/data/surrogate_cloud.npz,
`python -m torch_amd.train_surrogate --workload cloud --max-subcycles 131072`)
on the operator-split cloud's conditions.

The cloud's cooling is not stiff at the dt Torch::calculateTimeStep allows:
every cell finishes within the prepass's exact subcycles there.  The network
takes the cells that need more, which appear at multiples of that dt:
* fresh cells of the training distribution (seeds 9000+, disjoint from the
  training seeds 5000+): temperatures within 1% median and 5% p99 of the FP64
  integrator, and the trained network cuts the p99 error of its cells at
  least five-fold against the coarse estimate alone (a zero network);
* the cloud grid itself (96^3, 40 operator-split steps) cooled at 100x and
  1000x the loop's dt: every cell past the exact subcycles is inside the
  envelope and goes to the network, within the same error bounds."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
WEIGHTS = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "torch_amd", "data",
                       "surrogate_cloud.npz")


@pytest.mark.skipif(not os.path.exists(WEIGHTS), reason="cloud surrogate weights not trained yet")
def test_cloud_surrogate_accuracy_and_gain(gpu):
    import torch

    from torch_amd import surrogate as sg
    from torch_amd.train_surrogate import workload_inputs

    cool = gpu.Cooling(gpu.thermo_params())
    mlp = sg.attach(cool, WEIGHTS, gate=0.01)
    inp = workload_inputs(1 << 18, seed=9000, workload="cloud")
    cnt = sg.subcycles(cool, inp)
    inp = inp[:, (cnt >= 0) & (cnt <= 131072)].contiguous()
    r, ref = sg.compare(cool, inp, return_fp64=True)
    print(r)
    assert r["mlp_queued"] > 0.2 * r["cells"], r
    assert r["mlp_accepted"] > 0.9 * r["mlp_queued"], r
    assert r["rel_T_err_mlp_cells"]["median"] <= 0.01 and r["rel_T_err_mlp_cells"]["p99"] <= 0.05, r
    assert r["rel_T_err_all"]["median"] <= 0.01 and r["rel_T_err_all"]["p99"] <= 0.05, r

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


@pytest.mark.skipif(not os.path.exists(WEIGHTS), reason="cloud surrogate weights not trained yet")
def test_cloud_grid_at_100x_dt_goes_to_the_network(gpu):
    from torch_amd import cloud
    from torch_amd import surrogate as sg

    p, g, cool, thermo = cloud.build(96, "mixed")
    mlp = sg.attach(cool, WEIGHTS, gate=0.01)
    cool.run_split(g, 40, surrogate=True, **thermo)
    g.synchronize()
    t, dt, nst = g.clock()
    assert nst == 40
    for mult in (100.0, 1000.0):
        inp = cloud.cooling_inputs(g, p, thermo["col_den"], thermo["F_FUV"], thermo["excluded"], mult * dt)
        r = sg.compare(cool, inp)
        print(mult, r)
        past_exact = r["cells"] - r["exact"]
        assert past_exact > 0, r
        assert r["mlp_queued"] == past_exact, r            # none outside the envelope
        assert r["mlp_accepted"] > 0.9 * past_exact, r
        assert r["rel_T_err_mlp_cells"]["median"] <= 0.01 and r["rel_T_err_mlp_cells"]["p99"] <= 0.05, r
    g.close()
    del mlp


@pytest.mark.skipif(not os.path.exists(WEIGHTS), reason="cloud surrogate weights not trained yet")
def test_grid_substep_forms_inputs_in_the_prepass(gpu):
    """A surrogate thermo sub-step on the grid forms each cell's inputs in the
    prepass from U (writing the gathered array only for the cells it queues);
    with the heating record on it gathers the whole array first.  Both give
    the same state bit for bit (mixed and FP64 grids), at the loop's dt and
    at 100x it, where cells go to the network and the FP64 fallback, and the
    thermo time step agrees too."""
    import torch

    from torch_amd import cloud
    from torch_amd import surrogate as sg

    for precision in ("mixed", "fp64"):
        states, dts = [], []
        for track in (False, True):
            p, g, cool, thermo = cloud.build(48, precision)
            if track:
                g.heating_track(True)
            mlp = sg.attach(cool, WEIGHTS, gate=0.01)
            cool.run_split(g, 6, surrogate=True, **thermo)
            g.synchronize()
            t, dt, _ = g.clock()
            dts.append(cool.grid_timestep(g, p.dt_max, **thermo))
            for mult in (1.0, 100.0):
                cool.grid_substep(g, mult * dt, surrogate=True, **thermo)
            g.stage_dt(0, 0.5 * dt)   # a hydro sub-step on the float Q the cooling wrote
            g.synchronize()
            states.append(g.get_conserved())
            g.close()
            del mlp
        assert dts[0] == dts[1], (precision, dts)
        assert np.array_equal(states[0], states[1]), precision
