"""The device-resident operator-split loop (torchgpu_run_split) against the
same Torch::fullStep sequence driven from the host (dt of both operators
read back, every sub-step a separate call) on a small cloud set-up: the two
must give bitwise identical states, since they launch the same kernels with
the same dt arithmetic (0.5 dt, dt, dt / 2.0 as Torch.cpp:406-415)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_device_loop_matches_host_loop(gpu, precision):
    from torch_amd import cloud

    n, steps = 24, 6
    p1, g1, c1, th1 = cloud.build(n, precision)
    p2, g2, c2, th2 = cloud.build(n, precision)
    loop = cloud.HostLoop(g1, c1, th1, p1.dt_max, surrogate=False)
    dts = [loop.step() for _ in range(steps)]
    g1.synchronize()
    c2.run_split(g2, steps, **th2)
    g2.synchronize()
    t, dt, nst = g2.clock()
    assert nst == steps
    assert dt == dts[-1]
    assert t == pytest.approx(sum(dts), rel=1e-14)
    U1, U2 = g1.get_conserved(), g2.get_conserved()
    assert np.isfinite(U1).all()
    np.testing.assert_array_equal(U1, U2)
    # the state moved: hydro and cooling both acted
    U0 = cloud.build(n, precision)[1].get_conserved()
    assert not np.array_equal(U0[1], U2[1])


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_device_loop_matches_host_loop_surrogate(gpu, precision):
    """The same with the surrogate cooling step: the device loop applies the
    cooling update inside the surrogate prepass, defers the coarse
    estimates, takes the hydro CFL maximum from the thermo time step kernel
    and, in the steps that open with the cooling sub-step, the pre-step
    rates that kernel computed on the same state; the host loop evaluates
    each of them in its own call.  Bitwise the same states and dt."""
    import os

    from torch_amd import cloud
    from torch_amd import surrogate as sg

    weights = os.path.join(os.path.dirname(os.path.abspath(cloud.__file__)), "data", "surrogate_cloud.npz")
    n, steps = 24, 6
    p1, g1, c1, th1 = cloud.build(n, precision)
    p2, g2, c2, th2 = cloud.build(n, precision)
    m1, m2 = sg.attach(c1, weights), sg.attach(c2, weights)  # noqa: F841 (kept alive with the handles)
    loop = cloud.HostLoop(g1, c1, th1, p1.dt_max, surrogate=True)
    dts = [loop.step() for _ in range(steps)]
    g1.synchronize()
    c2.run_split(g2, steps, surrogate=True, **th2)
    g2.synchronize()
    t, dt, nst = g2.clock()
    assert nst == steps
    assert dt == dts[-1]
    assert t == pytest.approx(sum(dts), rel=1e-14)
    U1, U2 = g1.get_conserved(), g2.get_conserved()
    assert np.isfinite(U1).all()
    np.testing.assert_array_equal(U1, U2)
