"""The oracle's star geometry, CausalWind cells, FUV flux and column-density
ray trace (oracle/torch_oracle.c) against tests/golden/raytrace.npz, which
the reference's own Grid, Radiation and Thermodynamics translation units
wrote (tests/golden/make_golden.py raytrace): bitwise.  Four grids (ragged,
star at a corner, off-centre with a wind radius of 2, star at the far
corner) of random clumpy media; the pre-step RATE of the star-on thermo
preTimeStep follows from the oracle's thermo rates with these inputs."""
import os

import numpy as np
import pytest

import oracle

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "raytrace.npz")
CASES = ["a", "b", "c", "d"]


def _case(f, name):
    shape = tuple(int(x) for x in f[f"{name}_shape"])
    star = tuple(int(x) for x in f[f"{name}_star"])
    S, mh, pi, dx = (float(x) for x in f[f"{name}_info"])
    return shape, star, int(f[f"{name}_wind_radius"]), f[f"{name}_Q"], f[f"{name}_out"], S, mh, pi, dx


@pytest.mark.parametrize("name", CASES)
def test_geometry_and_ray_trace_match_reference(name):
    f = np.load(GOLD)
    shape, star, wind_r, Q, o, S, mh, pi, dx = _case(f, name)
    geom = oracle.star_geometry(shape, star, dx, pi)
    np.testing.assert_array_equal(geom[0], o[..., 0:4].astype(np.int64))
    np.testing.assert_array_equal(geom[1], o[..., 4:8])
    np.testing.assert_array_equal(geom[2], o[..., 8])
    np.testing.assert_array_equal(geom[3], o[..., 9])
    np.testing.assert_array_equal(oracle.star_wind(shape, star, wind_r), o[..., 14] > 0)
    col, dcol = oracle.thermo_ray_trace(shape, star, geom, mh, Q[0])
    np.testing.assert_array_equal(col, o[..., 10])
    np.testing.assert_array_equal(dcol, o[..., 11])
    # the column is nonzero away from the star's 3x3x3 block
    assert (col != 0).sum() > 0.5 * col.size


@pytest.mark.parametrize("name", CASES)
def test_star_on_pre_step_rates_match_reference(name):
    """RATE and HEAT of Thermodynamics::preTimeStepCalculations with the
    star's F_FUV and the ray-traced COL_DEN (Thermodynamics.cpp:218-262);
    CausalWind cells are skipped (RATE stays 0)."""
    f = np.load(GOLD)
    shape, star, wind_r, Q, o, S, mh, pi, dx = _case(f, name)
    n = Q[0].size
    geom = oracle.star_geometry(shape, star, dx, pi)
    col, _ = oracle.thermo_ray_trace(shape, star, geom, mh, Q[0])
    F = oracle.star_fuv(shape, star, dx, pi, S)
    wind = oracle.star_wind(shape, star, wind_r)
    c = oracle.converter(1e-20, 1e-28, 1e11)
    R = oracle.to_code(c, 8.314462e7, 0, 2, -2)
    dfloor = oracle.to_code(c, 1e-30, 1, -3, 0)
    t = oracle.thermo_params(c, pfloor=0.1 * R * dfloor * 0.1)
    inp = np.zeros((n, 9))
    inp[:, 0], inp[:, 1], inp[:, 2] = Q[0].ravel(), Q[1].ravel(), Q[2].ravel()
    inp[:, 3] = np.where(wind.ravel(), -1.0, Q[3].ravel())
    inp[:, 4] = Q[1].ravel() / 0.67
    inp[:, 5] = 100.0
    inp[:, 6] = col.ravel()
    inp[:, 7] = F.ravel()
    out = oracle.thermo_batch(t, inp, nthreads=4)
    live = inp[:, 3] >= 1e-2
    np.testing.assert_array_equal(out[:, 1], o[..., 12].ravel())
    np.testing.assert_array_equal(out[live, 0], o[..., 13].ravel()[live])
    assert live.sum() > 0 and (out[~live, 1] == 0).all()
