"""GPU subcycled cooling step against the reference fixtures and the oracle.

The GPU uses the device math library (ocml) for exp/log/log10/pow/sqrt, the
reference glibc's; agreement is therefore to a tolerance, stated per test."""
import os

import numpy as np
import pytest

from torch_amd import cooling_inputs

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _cooling(gpu, pfloor=None, **kw):
    p = gpu.thermo_params(**kw)
    if pfloor is not None:
        p.pfloor = pfloor
    return gpu.Cooling(p)


def test_thermo_params_match_oracle(gpu, orc):
    c = orc.converter(1e-20, 1e-28, 1e11)
    t = orc.thermo_params(c)
    p = gpu.thermo_params()
    for k in ("hydrogen_mass", "spec_gas_const", "dust_xsec", "imlc", "nmlc", "ciec", "n0", "nmc", "fuvh_a",
              "fuvh_b", "fuvh_c", "irh_a", "irh_b", "crh", "spl_min_slope", "spl_max_slope"):
        assert getattr(p, k) == getattr(t, k), k
    assert list(p.spl_y2) == list(t.spl_y2)


def test_thermo_cells_vs_reference_fixture(gpu):
    import torch

    f = np.load(os.path.join(GOLD, "thermo.npz"))
    inp = torch.from_numpy(np.ascontiguousarray(f["inputs"].T)).cuda()
    cool = _cooling(gpu, pfloor=float(f["pfloor"]))
    out, tot = cool.thermo_cells(inp)
    out = out.cpu().numpy()
    # heating and pre-step rate: one evaluation -> a few ulp of libm
    np.testing.assert_allclose(out[1], f["rate_pre"], rtol=1e-12, atol=0)
    ok = f["inputs"][:, 3] >= 1e-2
    np.testing.assert_allclose(out[0][ok], f["heat"][ok], rtol=1e-12, atol=0)
    # post-subcycle rate: after up to 2e5 explicit substeps, relative 1e-9
    rel = np.abs(out[2] - f["rate"]) / np.maximum(np.abs(f["rate"]), 1e-300)
    assert np.median(rel) < 1e-13
    assert np.quantile(rel, 0.999) < 1e-8, np.quantile(rel, [0.5, 0.99, 0.999, 1.0])


def test_thermo_cells_vs_oracle_workload_sample(gpu, orc):
    """Cells of the cooling workload (log-uniform rho, T, F, dt; uniform x,
    tau), including long subcycles; the subcycled pressure and the number of
    subcycles match the oracle."""
    import torch

    c = orc.converter(1e-20, 1e-28, 1e11)
    tc = lambda v, m, l, t: orc.to_code(c, v, m, l, t)  # noqa: E731
    s = cooling_inputs.sample_cgs(4096, seed=99)
    inp = cooling_inputs.to_code_inputs(s, tc)
    pfloor = 0.1 * tc(8.314462e7, 0, 2, -2) * tc(1e-30, 1, -3, 0) * 0.1
    t = orc.thermo_params(c, pfloor=pfloor)
    ref = orc.thermo_batch(t, inp, nthreads=8)
    cool = _cooling(gpu)
    assert cool.params.pfloor == pfloor
    out, tot = cool.thermo_cells(torch.from_numpy(np.ascontiguousarray(inp.T)).cuda())
    out = out.cpu().numpy()
    np.testing.assert_array_equal(out[3], ref[:, 3])          # subcycle counts
    assert tot == int(ref[:, 3].sum())
    rel = np.abs(out[4] - ref[:, 4]) / np.abs(ref[:, 4])      # subcycled pressure
    assert np.median(rel) < 1e-14 and rel.max() < 1e-6, np.quantile(rel, [0.5, 0.99, 1.0])


def test_thermo_gate_and_no_subcycling(gpu, orc):
    import torch

    c = orc.converter(1e-20, 1e-28, 1e11)
    tc = lambda v, m, l, t: orc.to_code(c, v, m, l, t)  # noqa: E731
    inp = cooling_inputs.to_code_inputs(cooling_inputs.sample_cgs(512, seed=3), tc)
    inp[:256, 3] = 0.0  # ADV below thermo_hii_switch: exact no-op
    cool = _cooling(gpu, subcycling=False)
    out, _ = cool.thermo_cells(torch.from_numpy(np.ascontiguousarray(inp.T)).cuda())
    out = out.cpu().numpy()
    assert (out[2][:256] == 0).all() and (out[3] == 0).all()
    np.testing.assert_array_equal(out[2][256:], out[1][256:])  # RATE unchanged by integrate


def test_subcycle_counts_match_oracle(gpu, orc):
    """torchgpu_thermo_subcycles gives, without integrating, the number of
    subcycles Thermodynamics::integrate takes (Thermodynamics.cpp:303-312):
    equal to the oracle's count on cooling-workload cells and on cells of the
    cloud's conditions (dense, unlit, dt to 1e13 s), -1 where dt / dti is
    past the int range, and no error is left on the handle."""
    import torch

    from torch_amd import surrogate as sg

    c = orc.converter(1e-20, 1e-28, 1e11)
    tc = lambda v, m, l, t: orc.to_code(c, v, m, l, t)  # noqa: E731
    pfloor = 0.1 * tc(8.314462e7, 0, 2, -2) * tc(1e-30, 1, -3, 0) * 0.1
    t = orc.thermo_params(c, pfloor=pfloor)
    cool = _cooling(gpu)
    inp = cooling_inputs.to_code_inputs(cooling_inputs.sample_cgs(4096, seed=99), tc)
    cnt = sg.subcycles(cool, torch.from_numpy(np.ascontiguousarray(inp.T)).cuda()).cpu().numpy()
    ref = orc.thermo_batch(t, inp, nthreads=8)
    np.testing.assert_array_equal(cnt, ref[:, 3])
    # the cloud's conditions; one cell with dt / dti past the int range
    cl = cooling_inputs.to_code_inputs(cooling_inputs.sample_cgs_cloud(4096, seed=7), tc)
    cl[0, 8] = 1e300
    cnt = sg.subcycles(cool, torch.from_numpy(np.ascontiguousarray(cl.T)).cuda()).cpu().numpy()
    assert cnt[0] == -1 and (cnt >= -1).all()
    keep = (cnt >= 0) & (cnt <= 100000)
    assert keep.sum() > 2500 and (cnt[keep] > 4).sum() > 100, (keep.sum(), (cnt[keep] > 4).sum())
    ref = orc.thermo_batch(t, np.ascontiguousarray(cl[keep]), nthreads=8)
    np.testing.assert_array_equal(cnt[keep], ref[:, 3])
    # the counting call raised nothing: the integrator runs on the kept cells
    out, tot = cool.thermo_cells(torch.from_numpy(np.ascontiguousarray(cl[keep].T)).cuda())
    np.testing.assert_array_equal(out[3].cpu().numpy(), ref[:, 3])
