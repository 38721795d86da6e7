"""Torch::calculateTimeStep's debug test on the device loops against the
reference harness's loop (tests/golden/debug_stop.npz): with Integration.debug
(the TorchParameters default, Parameters.hpp:50) Torch::run ends after the step
whose hydro dt has 100 dt / tmax <= 1e-6 (Torch.cpp:321-332).  A 16^3
converging flow, FP64: the step count, every dt, the time and the state are
bitwise the reference's, for a stop after step 2, a stop after a later step,
debug off (every step runs) and a tmax for which the test never fires."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _fixture():
    return np.load(os.path.join(GOLD, "debug_stop.npz"))


def _grid(gpu, f, case, precision=None):
    Q = f["Q"]
    tfloor = float(f["tfloor"])
    tmax = float(f[f"{case}_tmax"])
    p = gpu.make_params(Q.shape[1:], tfloor=tfloor, pfloor=0.1 * 8.314462e7 * 1e-30 * tfloor, tmax=tmax,
                        debug=bool(f[f"{case}_debug"]), precision=gpu.FP64 if precision is None else precision)
    g = gpu.Grid(p)
    g.set_primitive(np.ascontiguousarray(Q))
    return g


@pytest.mark.parametrize("case", ["early", "late", "nodebug", "run"])
def test_device_loop_debug_stop_matches_reference(gpu, case):
    f = _fixture()
    g = _grid(gpu, f, case)
    nsteps = int(f["nsteps"])
    # one step per call: a call after the stop does nothing
    dts = []
    for _ in range(nsteps):
        g.run(1)
        t, dt, n = g.clock()
        if n > len(dts):
            dts.append(dt)
    g.synchronize()
    t, dt, n = g.clock()
    assert n == int(f[f"{case}_steps"])
    assert g.quitting() == bool(f[f"{case}_quitting"])
    np.testing.assert_array_equal(np.array(dts), f[f"{case}_dts"])
    assert t == float(f[f"{case}_time"])
    np.testing.assert_array_equal(g.get_conserved(), f[f"{case}_U"])
    # one call for the whole run stops at the same step
    g2 = _grid(gpu, f, case)
    g2.run(nsteps)
    g2.synchronize()
    assert g2.clock() == (t, dt, n)
    np.testing.assert_array_equal(g2.get_conserved(), f[f"{case}_U"])
    # reset_clock clears the stop
    g2.reset_clock()
    assert not g2.quitting()


def test_mixed_path_debug_stop(gpu):
    """The mixed-precision path takes its CFL dt from the FP64-recovered
    primitives too: it stops after the same step."""
    f = _fixture()
    for case in ("early", "late"):
        g = _grid(gpu, f, case, precision=gpu.MIXED)
        g.run(int(f["nsteps"]))
        g.synchronize()
        assert g.clock()[2] == int(f[f"{case}_steps"]) and g.quitting(), case


@pytest.mark.parametrize("debug", [True, False])
def test_split_loop_debug_stop(gpu, debug):
    """torchgpu_run_split on the cloud set-up with a simulation time so long
    that the first CFL step already trips the test: it quits after step 2
    with debug on and runs on without."""
    from torch_amd import cloud

    p, g, cool, thermo = cloud.build(16, "fp64", tmax_s=1.0e24)
    if not debug:
        p.debug = 0
        g = gpu.Grid(p)
        g.set_primitive(np.ascontiguousarray(cloud.setup(16, lambda v, m, l, t: gpu.units(cloud.SCALES, v, m, l, t))[0]))
    cool.run_split(g, 5, **thermo)
    g.synchronize()
    t, dt, n = g.clock()
    assert g.quitting() == debug
    assert n == (2 if debug else 5)
    assert 100.0 * dt / p.tmax <= 1e-6


def test_run_printing_ends_at_stop(gpu, tmp_path):
    """Torch::run with its snapshots ends at the debug stop as the plain loop
    does: the initial print, no checkpoint reached, and no final print
    (isFinalPrintOn is only set by a print inside the loop, Torch.cpp:270-297)."""
    f = _fixture()
    g = _grid(gpu, f, "late")
    g2 = _grid(gpu, f, "late")
    files = gpu.run_printing(g, int(f["nsteps"]), str(tmp_path), scales=(1.0, 1.0, 1.0))
    g.synchronize()
    assert g.quitting() and g.clock()[2] == int(f["late_steps"])
    # the initial print and no checkpoint inside the run (ncheckpoints = 1):
    # no final print is due
    assert [os.path.basename(x) for x in files] == ["data2D_000000.txt.gz"]
    g2.run(int(f["nsteps"]))
    g2.synchronize()
    np.testing.assert_array_equal(g.get_conserved(), g2.get_conserved())
