"""Torch::fullStep with hydrodynamics and thermodynamics (star off) on the
GPU against the reference's own units driven by the harness's loop
(tests/golden/cloud_split.npz): a 16^3 version of the operator-split cloud
(side 0.5 pc, the softened profile about a centre 0.35 pc from the grid
centre, T = 300 K, ionised core, ADV = 1) at the reference's unit scales,
6 steps.  The first step is dt_max 1e-20, then min(hydro CFL, thermo
|E/RATE|); the Strang sequence alternates T(dt/2) H(dt) T(dt/2) and
H(dt/2) T(dt) H(dt/2) with single-stage hydro sub-steps (Torch.cpp:389-418).
With the star off the reference's cooling sees no FUV flux and no column
density, and skips the star's cell (CausalWind, Fluid.cpp:41-54) at the grid
centre.

Bounds: the cooling kernels use the device exp / log / pow, so the FP64 path
is compared to rounding (every dt within 1e-10 relative, U within 1e-8
relative L-inf per variable).  The surrogate step integrates every cell with
at most 4 subcycles exactly in FP64 and sends the rest to the network (1 %
gate, FP64 fallback); with FP64 hydro it is held to the same bounds as the
integrator when the network takes no cell, else to 1e-3 relative L1 of the
energy.  The mixed-precision hydro (the cloud workload's configuration) is
held to 1e-6 relative L1 of density and 1e-5 of energy and momenta."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
SCALES = (1e-20, 1e-28, 1e11)
WEIGHTS = os.path.join(os.path.dirname(GOLD), "..", "torch_amd", "data", "surrogate_cooling.npz")


def _cloud(gpu, precision):
    f = np.load(os.path.join(GOLD, "cloud_split.npz"))
    Q = f["Q"]
    n = Q.shape[1]
    to = np.vectorize(lambda v, m, l, t: gpu.units(SCALES, v, m, l, t))
    R = gpu.units(SCALES, 8.314462e7, 0, 2, -2)
    dfloor = gpu.units(SCALES, 1e-30, 1, -3, 0)
    p = gpu.make_params((n, n, n), side_length=gpu.units(SCALES, float(f["side"]), 0, 1, 0), gamma=1.67,
                        dfloor=dfloor, pfloor=0.1 * R * dfloor * 0.1, tfloor=0.1, spec_gas_const=R,
                        tmax=gpu.units(SCALES, 1e13, 0, 0, 1), ncheckpoints=10,
                        precision=gpu.MIXED if precision == "mixed" else gpu.FP64)
    # Torch::toCodeUnits, value by value
    Qc = Q.copy()
    Qc[0] = to(Q[0], 1, -3, 0)
    Qc[1] = to(Q[1], 1, -1, -2)
    for a in range(3):
        Qc[4 + a] = to(Q[4 + a], 0, 1, -1)
    g = gpu.Grid(p)
    g.set_primitive(np.ascontiguousarray(Qc))
    tp = gpu.thermo_params(scales=SCALES, thermo_hii_switch=1e-2, temperature_floor=0.1)
    tp.pfloor = p.pfloor
    cool = gpu.Cooling(tp)
    c = n // 2
    return f, g, cool, c + n * (c + n * c)


def _rel_linf(U, Ur):
    out = []
    for v in range(7):
        s = np.abs(Ur[v]).max()
        out.append(0.0 if s == 0 else np.abs(U[v] - Ur[v]).max() / s)
    return np.array(out)


def _rel_l1(U, Ur):
    return np.array([np.abs(U[v] - Ur[v]).sum() / max(np.abs(Ur[v]).sum(), 1e-300) for v in range(7)])


def _run(gpu, precision, surrogate, host_loop=False):
    f, g, cool, excl = _cloud(gpu, precision)
    mlp = None
    if surrogate:
        from torch_amd import surrogate as sg

        mlp = sg.attach(cool, WEIGHTS)
    nsteps = len(f["dts"])
    dts, states = [], []
    if host_loop:
        from torch_amd import cloud

        loop = cloud.HostLoop(g, cool, {"excluded": excl}, g.params.dt_max, surrogate)
        for _ in range(nsteps):
            dts.append(loop.step())
            g.synchronize()
            states.append(g.get_conserved())
    else:
        for s in range(1, nsteps + 1):
            cool.run_split(g, 1, excluded=excl, surrogate=surrogate)
            t, dt, n = g.clock()
            assert n == s
            dts.append(dt)
            g.synchronize()
            states.append(g.get_conserved())
    return f, np.array(dts), states, mlp


@pytest.mark.parametrize("host_loop", [False, True])
def test_cloud_split_fp64_matches_reference(gpu, host_loop):
    f, dts, states, _ = _run(gpu, "fp64", False, host_loop)
    np.testing.assert_allclose(dts, f["dts"], rtol=1e-10, atol=0)
    for s, U in enumerate(states, 1):
        rel = _rel_linf(U, f[f"U_{s}"])
        assert (rel < 1e-8).all(), (s, rel)
    # cooling acted: the energy moved away from pure hydro's
    assert not np.array_equal(states[-1][1], states[0][1])


def test_cloud_split_surrogate_matches_reference(gpu):
    f, dts, states, mlp = _run(gpu, "fp64", True)
    rel = _rel_linf(states[-1], f[f"U_{len(dts)}"])
    l1 = _rel_l1(states[-1], f[f"U_{len(dts)}"])
    print("surrogate: dts rel", np.abs(dts / f["dts"] - 1).max(), "U rel Linf", rel, "rel L1", l1)
    np.testing.assert_allclose(dts, f["dts"], rtol=1e-6, atol=0)
    assert l1[1] < 1e-3, l1
    assert (l1[[0, 4, 5, 6]] < 1e-6).all(), l1


def test_cloud_split_mixed_surrogate_matches_reference(gpu):
    """The cloud workload's configuration: mixed-precision hydro sub-steps and
    the surrogate cooling step, against the reference's FP64 fullStep."""
    f, dts, states, mlp = _run(gpu, "mixed", True)
    U, Ur = states[-1], f[f"U_{len(dts)}"]
    l1 = _rel_l1(U, Ur)
    print("mixed+surrogate: dts rel", np.abs(dts / f["dts"] - 1).max(), "rel L1", l1, "rel Linf", _rel_linf(U, Ur))
    np.testing.assert_allclose(dts, f["dts"], rtol=1e-5, atol=0)
    assert l1[0] < 1e-6, l1
    assert (l1[[1, 4, 5, 6]] < 1e-5).all(), l1


def test_thermo_nsteps_overflow_reported(gpu):
    """A thermo sub-step whose dt needs more subcycles than an int holds: the
    kernel marks the overflow and returns (no subcycling loop runs), the
    grid's next synchronisation reports TORCHGPU_ERR_NSTEPS_RANGE, and the
    cooling handle's error word is cleared once it has been reported."""
    f, g, cool, excl = _cloud(gpu, "fp64")
    cool.grid_substep(g, 1.0e30, excluded=excl)
    with pytest.raises(gpu.TorchGpuError) as e:
        g.synchronize()
    assert e.value.code == 9, e.value
    # the handle itself works on: a fresh grid's sub-step at a normal dt
    f, g2, _, _ = _cloud(gpu, "fp64")
    cool.grid_substep(g2, float(f["dts"][1]), excluded=excl)
    g2.synchronize()
