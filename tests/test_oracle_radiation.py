"""The oracle's Radiation integrator (oracle/torch_oracle.c: initField,
preTimeStepCalculations, calculateTimeStep with its first-call quirk, the
causal transfer with the implicit / implicit2 / explicit HII update, the
tti / neq / off couplings) against tests/golden/radiation.npz, which the
reference's own Radiation, Grid and Fluid units wrote
(tests/golden/make_golden.py radiation): bitwise, every dt and the final
U, R and Q[HII, ADV]."""
import os
import sys

import numpy as np
import pytest

import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
from make_golden import RAD_CASES  # noqa: E402

GOLD = os.path.join(HERE, "golden", "radiation.npz")
R_CGS = 8.314462e7


def rad_setup(name):
    cs = RAD_CASES[name]
    c = oracle.converter(1e-20, 1e-28, 1e11)
    tc = lambda v, m, l, t: oracle.to_code(c, v, m, l, t)  # noqa: E731
    shape = cs["shape"]
    kv = cs["kv"]
    R = tc(R_CGS, 0, 2, -2)
    dfloor = tc(1e-30, 1, -3, 0)
    p = oracle.make_params(shape, tc(cs["side"], 0, 1, 0) / shape[2], gamma=1.67, dfloor=dfloor,
                           pfloor=0.1 * R * dfloor * 0.1, tfloor=0.1, spec_gas_const=R,
                           dt_max=tc(1e13, 0, 0, 1) / 10.0, nthreads=1)
    r = oracle.rad_params(c, K=(kv.get("K1", 0.0), kv.get("K2", 0.0), kv.get("K3", 0.0), kv.get("K4", 0.0)),
                          sigma=6.3e-18, alphaB=2.59e-13, tau0=0.6, THI=300, THII=1e4,
                          scheme=kv["integration_scheme"], coupling=kv["coupling"],
                          collisions_on=bool(kv["collisions_on"]),
                          heating_amplification=kv.get("heating_amplification", 1.0), photon_energy=2.976e-11,
                          photon_rate=1e49)
    return cs, p, r


@pytest.mark.parametrize("name", list(RAD_CASES))
def test_radiation_matches_reference(name):
    f = np.load(GOLD)
    cs, p, r = rad_setup(name)
    Q = f[f"{name}_Q"]
    Qf = oracle.fix_primitives(p, Q, np.zeros_like(Q))
    U0 = oracle.UfromQ_grid(Qf, p.gamma)
    R0 = oracle.rad_init_field(r, cs["shape"], cs["star"], p.dx, Qf)
    U, R, q, dts, _ = oracle.rad_run(r, p, cs["star"], cs["wind"], U0, R0, cs["nsteps"])
    np.testing.assert_array_equal(dts, f[f"{name}_dts"])
    np.testing.assert_array_equal(U, f[f"{name}_U"])
    np.testing.assert_array_equal(R, f[f"{name}_R"])
    np.testing.assert_array_equal(q, f[f"{name}_q"])
    # the transfer did something: HII changed and the front cells heat
    assert not np.array_equal(U[2], U0[2])


def test_first_nonwind_cell_follows_causal_order():
    """The first cell of the reference's CausalNonWind list (buildCausal's
    octant snake): next to the star along +x unless that is a wind cell or
    off the grid."""
    assert oracle.first_nonwind_cell((8, 9, 10), (3, 4, 2), 1) == 224  # (4, 4, 2)
    assert oracle.first_nonwind_cell((8, 9, 10), (9, 4, 2), 0) == 9 + 10 * (5 + 9 * 2)  # x edge: y + 1 row
    # wind radius 2: (sx + 2) has d2 = 2 * 2.5 = 5 > 4
    assert oracle.first_nonwind_cell((8, 9, 10), (3, 4, 2), 2) == 5 + 10 * (4 + 9 * 2)


def test_recombination_spline_ends():
    """LinearSplineData: the Hummer table at its knots, power-law ends with
    the end slopes."""
    c = oracle.converter(1.0, 1.0, 1.0)
    r = oracle.rad_params(c)
    rec = lambda T: oracle.lib().orc_rad_recombination(oracle.ctypes.byref(r), T)  # noqa: E731
    T0 = np.exp(np.log(10.0) * 1.0)
    assert rec(T0) == pytest.approx(9.283e-11 / np.sqrt(T0), rel=1e-15)
    T30 = np.exp(np.log(10.0) * 7.0)
    assert rec(T30) == pytest.approx(3.498e-13 / np.sqrt(T30), rel=1e-15)
    s = (np.log10(3.498e-13 / np.sqrt(T30)) - np.log10(5.199e-13 / np.sqrt(T30 / 10 ** 0.2))) / 0.2
    assert rec(2 * T30) == pytest.approx(3.498e-13 / np.sqrt(T30) * 2.0 ** s, rel=1e-12)
