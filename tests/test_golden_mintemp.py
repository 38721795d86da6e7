"""The min_temp_initial_state fixture (tests/golden/mintemp.npz, written by
make_golden.mintemp through the reference's own units): its inputs are the
generator's, and the option changes the run -- the reference's T_min field
(cgs p / rho over the code-unit gas constant, ~1e-8 of the gas temperature
at the standard scales) removes the soft landing that the constant 100 K
applies to this 150-250 K gas, so the runs part from the second step."""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
from make_golden import MINTEMP, mintemp_q  # noqa: E402


def test_mintemp_fixture_inputs_and_effect():
    f = np.load(os.path.join(HERE, "golden", "mintemp.npz"))
    np.testing.assert_array_equal(f["Q"], mintemp_q())
    cs = MINTEMP
    assert f["dts_on"].shape == (cs["nsteps"],) and f["dts_off"].shape == (cs["nsteps"],)
    # the first step is dt_max * 1e-20 either way, the second already differs
    assert f["dts_on"][0] == f["dts_off"][0]
    assert f["dts_on"][1] != f["dts_off"][1]
    # without the 100 K soft landing the gas cools further
    assert f["U_on"][1].sum() < f["U_off"][1].sum()
