"""The oracle's Thermodynamics::fillHeatingArrays (oracle/torch_oracle.c,
orc_fill_heating) against tests/golden/heating.npz, which the reference's
own translation units wrote through the harness's Torch::run loop
(tests/golden/make_golden.py heating): starting from the GridCell::H the
loop left, two fills on Fluid::Q give the reference's H bitwise -- the eight
terms recomputed from Q, RHII and EUVH as the radiation pre-step set them,
and TOT accumulating RHII + EUVH over the fills.  Cases: a star with
cooling and radiation, a star with radiation only (TOT never reset by an
integrate) and the star-off cooling cloud.  A fill that overwrote TOT
instead of adding to it, or that dropped the star's ray trace or FUV flux,
fails the second-fill or the FUVH/IRH columns."""
import os
import sys

import numpy as np
import pytest

import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
from make_golden import CLOUD_SPLIT, STAR_FULLSTEP  # noqa: E402

SCALES = (1e-20, 1e-28, 1e11)


def _inputs(f, name):
    c = oracle.converter(*SCALES)
    t = oracle.thermo_params(c)
    Q = f[f"{name}_Qprint"]
    shape = Q.shape[1:]
    n = shape[0]
    if name in ("star", "radonly"):
        cs = STAR_FULLSTEP
        dx = oracle.to_code(c, cs["side"], 0, 1, 0) / n
        geom = oracle.star_geometry(shape, cs["star"], dx, t.pi)
        col, _ = oracle.thermo_ray_trace(shape, cs["star"], geom, t.hydrogen_mass, Q[0])
        F = oracle.star_fuv(shape, cs["star"], dx, t.pi, oracle.to_code(c, cs["kv"]["photon_rate"], 0, 0, -1))
        wind = oracle.star_wind(shape, cs["star"], cs["wind"])
    else:
        col, F = None, None
        wind = np.zeros(shape, dtype=bool)
        m = CLOUD_SPLIT["n"] // 2
        wind[m, m, m] = True
    return t, Q, col, F, wind


@pytest.mark.parametrize("name", ["star", "radonly", "cloud"])
def test_fill_heating_matches_reference(name):
    f = np.load(os.path.join(HERE, "golden", "heating.npz"))
    t, Q, col, F, wind = _inputs(f, name)
    H0 = f[f"{name}_H0"]
    heat = np.ascontiguousarray(H0[[10, 2, 6]])
    for k in (1, 2):
        H = oracle.fill_heating(t, Q, heat, col_den=col, F=F, wind=wind)
        np.testing.assert_array_equal(H, f[f"{name}_H{k}"], err_msg=f"{name} fill {k}")
    if name == "radonly":
        # without an integrate between the fills TOT grows by RHII + EUVH
        live = ~wind & (Q[3] >= 1e-2)
        d = f["radonly_H2"][10] - f["radonly_H1"][10]
        np.testing.assert_allclose(d[live], (H0[2] + H0[6])[live], rtol=1e-12, atol=0)
        assert (np.abs(H0[2] + H0[6])[live] > 0).any()
    if name != "cloud":
        assert (f[f"{name}_H1"][7] != 0).any()  # FUV heating from the star
