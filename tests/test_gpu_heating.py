"""Thermodynamics::fillHeatingArrays on the GPU (torchgpu_fill_heating with
the grid's kept GridCell::H terms, torchgpu_heating_track) against
tests/golden/heating.npz, which the reference's own translation units wrote
through the harness's Torch::run loop: after the run, two fills on the
printed Fluid::Q.  Cases: a star with cooling and radiation (device loop of
all three components), the star with radiation only (TOT accumulates
RHII + EUVH over the fills, no integrate resets it) and the star-off
cooling cloud.  The GPU's cooling and radiation use the device exp / log /
pow and the terms are exponentials of the temperature, so each of the 11
columns is held to 1e-6 of its largest magnitude (the printed state to 1e-8
relative L-inf per variable, as the composed step's state); the zero pattern
(wind cells, cells below the HII switch) is exact."""
import os
import sys

import numpy as np
import pytest

import oracle as orc

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
from make_golden import CLOUD_SPLIT, STAR_FULLSTEP  # noqa: E402

SCALES = (1e-20, 1e-28, 1e11)


def _fixture():
    return np.load(os.path.join(HERE, "golden", "heating.npz"))


def _grid(gpu, Q, side, tmax=1e13):
    n = Q.shape[1]
    to = np.vectorize(lambda v, m, l, t: gpu.units(SCALES, v, m, l, t))
    R = gpu.units(SCALES, 8.314462e7, 0, 2, -2)
    dfloor = gpu.units(SCALES, 1e-30, 1, -3, 0)
    p = gpu.make_params((n, n, n), side_length=gpu.units(SCALES, side, 0, 1, 0), gamma=1.67, dfloor=dfloor,
                        pfloor=0.1 * R * dfloor * 0.1, tfloor=0.1, spec_gas_const=R,
                        tmax=gpu.units(SCALES, tmax, 0, 0, 1), ncheckpoints=10)
    Qc = Q.copy()
    Qc[0] = to(Q[0], 1, -3, 0)
    Qc[1] = to(Q[1], 1, -1, -2)
    for a in range(3):
        Qc[4 + a] = to(Q[4 + a], 0, 1, -1)
    g = gpu.Grid(p)
    g.heating_track(True)
    g.set_primitive(np.ascontiguousarray(Qc))
    tp = gpu.thermo_params(scales=SCALES, thermo_hii_switch=1e-2, temperature_floor=0.1)
    tp.pfloor = p.pfloor
    return g, gpu.Cooling(tp), Qc


def _run(gpu, name, f):
    nsteps = int(f[f"{name}_nsteps"])
    if name == "cloud":
        g, cool, _ = _grid(gpu, f["cloud_Q0"], float(np.load(os.path.join(HERE, "golden", "cloud_split.npz"))["side"]))
        m = CLOUD_SPLIT["n"] // 2
        excl = m + CLOUD_SPLIT["n"] * (m + CLOUD_SPLIT["n"] * m)
        g.pause_at_checkpoints(True)
        cool.run_split(g, nsteps, excluded=excl)
        return g, cool, {"excluded": excl}
    cs = STAR_FULLSTEP
    kv = cs["kv"]
    g, cool, Qc = _grid(gpu, f[f"{name}_Q0"], cs["side"])
    star = gpu.make_star(cs["star"], cs["wind"], gpu.units(SCALES, kv["photon_rate"], 0, 0, -1))
    cool.set_star(g, star)
    rad = gpu.Radiation(gpu.rad_params(SCALES, K=(kv["K1"], 0, 0, 0), coupling=kv["coupling"],
                                       scheme=kv["integration_scheme"], photon_energy=kv["photon_energy"]), g, star)
    rad.init_field(Qc)
    g.pause_at_checkpoints(True)
    rad.run_full(nsteps, cooling=cool if name == "star" else None)
    g._rad = rad
    return g, cool, {}


@pytest.mark.parametrize("name", ["star", "radonly", "cloud"])
def test_fill_heating_matches_reference(gpu, name):
    f = _fixture()
    g, cool, kw = _run(gpu, name, f)
    g.synchronize()
    t, dt, n = g.clock()
    assert n == int(f[f"{name}_nsteps"])
    np.testing.assert_allclose(t, float(f[f"{name}_time"]), rtol=1e-10)
    Qp = g.print_state()
    Qr = f[f"{name}_Qprint"]
    for v in range(7):
        s = np.abs(Qr[v]).max()
        assert s == 0 or np.abs(Qp[v] - Qr[v]).max() / s < 1e-8, (name, v)
    for k in (1, 2):
        H = cool.fill_heating(g, **kw)
        Hr = f[f"{name}_H{k}"]
        np.testing.assert_array_equal(H == 0, Hr == 0, err_msg=f"{name} fill {k}: zero pattern")
        # RHII and EUVH are x (soft / rate), NaN where the radiation rate is 0
        # (Radiation.cpp:278-279), and TOT then too
        np.testing.assert_array_equal(np.isnan(H), np.isnan(Hr), err_msg=f"{name} fill {k}: NaN pattern")
        errs = []
        for i in range(11):
            ok = np.isfinite(Hr[i])
            s = np.abs(Hr[i][ok]).max() if ok.any() else 0.0
            errs.append(0.0 if s == 0 else float(np.abs(H[i][ok] - Hr[i][ok]).max() / s))
        print(name, k, "max rel error per column", ["%.1e" % e for e in errs])
        assert max(errs) < 1e-6, (name, k, errs)
