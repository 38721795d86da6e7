"""Thermodynamics.min_temp_initial_state against the reference's own units
(tests/golden/mintemp.npz: the harness's Torch driver loop with hydro and
thermo, 12^3 cool 150-250 K gas, 6 steps, the option on and off).  With the
option on, Thermodynamics::initialiseMinTempField sets each cell's T_min to
Fluid::calcTemperature of the cgs set-up with the code-unit gas constant,
before Torch::toCodeUnits (Torch.cpp:113-117, Thermodynamics.cpp:102-111);
off, T_min is 100 K.  The field goes to the cooling handle
(Cooling.set_min_temp) and the run is driven by FullStep (host loop) and by
the device loop.  The cooling kernels use the device exp / log / pow, so
the comparison is to rounding: every dt within 1e-10 relative, U with
relative L-inf below 1e-8 per variable."""
import os
import sys

import numpy as np
import pytest

import oracle as orc

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
from make_golden import MINTEMP  # noqa: E402

SCALES = (1e-20, 1e-28, 1e11)


def _close_U(U, Ur, what):
    for v in range(7):
        scale = np.abs(Ur[v]).max()
        if scale == 0:
            assert np.abs(U[v]).max() == 0, (what, v)
            continue
        rel = np.abs(U[v] - Ur[v]).max() / scale
        assert rel < 1e-8, (what, v, rel)


@pytest.mark.parametrize("device_loop", [False, True])
@pytest.mark.parametrize("on", [True, False])
def test_min_temp_initial_state_matches_reference(gpu, on, device_loop):
    from torch_amd import fullstep

    f = np.load(os.path.join(HERE, "golden", "mintemp.npz"))
    cs = MINTEMP
    n = cs["n"]
    c = orc.converter(*SCALES)
    tc = lambda v, m, l, t: orc.to_code(c, v, m, l, t)  # noqa: E731
    R = tc(8.314462e7, 0, 2, -2)
    dfloor = tc(1e-30, 1, -3, 0)
    p = gpu.make_params((n, n, n), side_length=tc(cs["side"], 0, 1, 0), gamma=1.67, dfloor=dfloor,
                        pfloor=0.1 * R * dfloor * 0.1, tfloor=0.1, spec_gas_const=R, tmax=tc(1e13, 0, 0, 1),
                        ncheckpoints=10)
    Q = f["Q"]
    Qc = Q.copy()
    to = np.vectorize(lambda v, m, l, t: gpu.units(SCALES, v, m, l, t))
    Qc[0] = to(Q[0], 1, -3, 0)
    Qc[1] = to(Q[1], 1, -1, -2)
    g = gpu.Grid(p)
    g.set_primitive(np.ascontiguousarray(Qc))
    tp = gpu.thermo_params(scales=SCALES, thermo_hii_switch=cs["kv"]["thermo_hii_switch"], temperature_floor=0.1)
    tp.pfloor = p.pfloor
    cool = gpu.Cooling(tp)
    if on:
        # Fluid::calcTemperature on the cgs primitives, X = 1
        mu_inv = 1.0 * (Q[2] + 1.0) + (1.0 - 1.0) * 0.25
        cool.set_min_temp((Q[1] / Q[0]) * (1.0 / mu_inv) / R)
    key = "on" if on else "off"
    # Thermodynamics walks CausalNonWind only: with the star off that skips
    # the cell at the default Star.cell_position (0, 0, 0) (Fluid.cpp:41-54)
    if device_loop:
        dts = []
        for _ in range(cs["nsteps"]):
            cool.run_split(g, 1, excluded=0)
            dts.append(g.clock()[1])
    else:
        loop = fullstep.FullStep(g, cooling=cool, thermo={"excluded": 0})
        dts = [loop.step() for _ in range(cs["nsteps"])]
    g.synchronize()
    np.testing.assert_allclose(np.array(dts), f[f"dts_{key}"], rtol=1e-10, atol=0)
    _close_U(g.get_conserved(), f[f"U_{key}"], key)
