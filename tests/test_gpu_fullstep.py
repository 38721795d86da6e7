"""Torch::fullStep with hydrodynamics, thermodynamics and radiation around a
star on the GPU (torch_amd.fullstep.FullStep over the library's hydro
sub-step, star-on cooling and Radiation calls) against the reference's own
units driven by the same loop (tests/golden/star_fullstep.npz, the harness's
Torch::fullStep with three components): 16^3 clumpy neutral medium, star
with wind radius 1, 4 steps.  The first step is dt_max 1e-20, then
min(hydro, radiation, thermo); the Strang rotation runs over three
components.  The cooling and radiation kernels use the device exp / log /
pow, so the comparison is to rounding: every dt within 1e-10 relative, U
with relative L-inf below 1e-8 per variable."""
import os
import sys

import numpy as np
import pytest

import oracle as orc

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
from make_golden import STAR_FULLSTEP  # noqa: E402

SCALES = (1e-20, 1e-28, 1e11)


def test_star_fullstep_matches_reference(gpu):
    from torch_amd import fullstep

    f = np.load(os.path.join(HERE, "golden", "star_fullstep.npz"))
    cs = STAR_FULLSTEP
    kv = cs["kv"]
    n = cs["n"]
    c = orc.converter(*SCALES)
    tc = lambda v, m, l, t: orc.to_code(c, v, m, l, t)  # noqa: E731
    R = tc(8.314462e7, 0, 2, -2)
    dfloor = tc(1e-30, 1, -3, 0)
    p = gpu.make_params((n, n, n), side_length=tc(cs["side"], 0, 1, 0), gamma=1.67, dfloor=dfloor,
                        pfloor=0.1 * R * dfloor * 0.1, tfloor=0.1, spec_gas_const=R, tmax=tc(1e13, 0, 0, 1),
                        ncheckpoints=10)
    # Torch::toCodeUnits of the cgs set-up, value by value
    Q = f["Q"]
    Qc = Q.copy()
    to = np.vectorize(lambda v, m, l, t: gpu.units(SCALES, v, m, l, t))
    Qc[0] = to(Q[0], 1, -3, 0)
    Qc[1] = to(Q[1], 1, -1, -2)
    for a in range(3):
        Qc[4 + a] = to(Q[4 + a], 0, 1, -1)
    g = gpu.Grid(p)
    g.set_primitive(np.ascontiguousarray(Qc))
    star = gpu.make_star(cs["star"], cs["wind"], gpu.units(SCALES, kv["photon_rate"], 0, 0, -1))
    tp = gpu.thermo_params(scales=SCALES, thermo_hii_switch=kv["thermo_hii_switch"], temperature_floor=0.1)
    tp.pfloor = p.pfloor
    cool = gpu.Cooling(tp)
    cool.set_star(g, star)
    rad = gpu.Radiation(gpu.rad_params(SCALES, K=(kv["K1"], 0, 0, 0), coupling=kv["coupling"],
                                       scheme=kv["integration_scheme"], photon_energy=kv["photon_energy"]), g, star)
    rad.init_field(Qc)
    loop = fullstep.FullStep(g, cooling=cool, radiation=rad)
    dts = []
    for step in range(1, cs["nsteps"] + 1):
        dts.append(loop.step())
        key = f"U_{step}"
        if key in f.files:
            g.synchronize()
            U, Ur = g.get_conserved(), f[key]
            for v in range(7):
                scale = np.abs(Ur[v]).max()
                if scale == 0:
                    assert np.abs(U[v]).max() == 0, (step, v)
                    continue
                rel = np.abs(U[v] - Ur[v]).max() / scale
                assert rel < 1e-8, (step, v, rel)
    np.testing.assert_allclose(np.array(dts), f["dts"], rtol=1e-10, atol=0)
    # the star ionised gas and heated it
    U0 = orc.UfromQ_grid(Qc, 1.67)
    assert (f[f"U_{cs['nsteps']}"][2] > U0[2] + 1e-30).any()
