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


def _star_run(gpu, cooling=True, tmax=None, dt_max=None):
    """Grid, cooling (or None), radiation, code-unit Q and hydro params of the
    fixture's set-up."""
    f = np.load(os.path.join(HERE, "golden", "star_fullstep.npz"))
    cs = STAR_FULLSTEP
    kv = cs["kv"]
    n = cs["n"]
    c = orc.converter(*SCALES)
    tc = lambda v, m, l, t: orc.to_code(c, v, m, l, t)  # noqa: E731
    R = tc(8.314462e7, 0, 2, -2)
    dfloor = tc(1e-30, 1, -3, 0)
    p = gpu.make_params((n, n, n), side_length=tc(cs["side"], 0, 1, 0), gamma=1.67, dfloor=dfloor,
                        pfloor=0.1 * R * dfloor * 0.1, tfloor=0.1, spec_gas_const=R,
                        tmax=tc(1e13, 0, 0, 1) if tmax is None else tmax, dt_max=dt_max, ncheckpoints=10)
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
    cool = gpu.Cooling(tp) if cooling else None
    if cool is not None:
        cool.set_star(g, star)
    rad = gpu.Radiation(gpu.rad_params(SCALES, K=(kv["K1"], 0, 0, 0), coupling=kv["coupling"],
                                       scheme=kv["integration_scheme"], photon_energy=kv["photon_energy"]), g, star)
    rad.init_field(Qc)
    return f, g, cool, rad, Qc, p


def _compare_U(g, f, step):
    key = f"U_{step}"
    if key not in f.files:
        return
    g.synchronize()
    U, Ur = g.get_conserved(), f[key]
    for v in range(7):
        scale = np.abs(Ur[v]).max()
        if scale == 0:
            assert np.abs(U[v]).max() == 0, (step, v)
            continue
        rel = np.abs(U[v] - Ur[v]).max() / scale
        assert rel < 1e-8, (step, v, rel)


def test_star_fullstep_matches_reference(gpu):
    from torch_amd import fullstep

    f, g, cool, rad, Qc, p = _star_run(gpu)
    cs = STAR_FULLSTEP
    loop = fullstep.FullStep(g, cooling=cool, radiation=rad)
    dts = []
    for step in range(1, cs["nsteps"] + 1):
        dts.append(loop.step())
        _compare_U(g, f, step)
    np.testing.assert_allclose(np.array(dts), f["dts"], rtol=1e-10, atol=0)
    # the star ionised gas and heated it
    U0 = orc.UfromQ_grid(Qc, 1.67)
    assert (f[f"U_{cs['nsteps']}"][2] > U0[2] + 1e-30).any()


def test_star_device_loop_matches_reference(gpu):
    """The same run with the whole Torch::run loop on the device
    (torchgpu_run_full: dt, the pre-step calculations and the three-component
    Strang sub-steps chosen and launched without a host round trip)."""
    f, g, cool, rad, Qc, p = _star_run(gpu)
    cs = STAR_FULLSTEP
    dts = []
    for step in range(1, cs["nsteps"] + 1):
        rad.run_full(1, cooling=cool)
        t, dt, n = g.clock()
        assert n == step
        dts.append(dt)
        _compare_U(g, f, step)
    np.testing.assert_allclose(np.array(dts), f["dts"], rtol=1e-10, atol=0)


@pytest.mark.parametrize("cooling", [True, False])
def test_star_device_loop_equals_host_loop(gpu, cooling):
    """torchgpu_run_full against FullStep's host-driven calls of the same
    kernels, bit for bit (with and without cooling), and the state printed at
    checkpoints: fixPrimitives(Q(U)) of the last sub-step's input, with the
    HII fraction the radiation sub-step leaves in Fluid::Q when it is the
    last (every third step)."""
    from torch_amd import fullstep

    nsteps = 6
    _, g1, cool1, rad1, _, p = _star_run(gpu, cooling)
    _, g2, cool2, rad2, _, _ = _star_run(gpu, cooling)
    hp = orc.make_params((p.nz, p.ny, p.nx), p.side_length / p.nx, gamma=p.gamma, dfloor=p.dfloor,
                         pfloor=p.pfloor, tfloor=p.tfloor, mass_frac_h=p.mass_fraction_h,
                         spec_gas_const=p.spec_gas_const)
    expected = {}

    class Recorder(fullstep.FullStep):
        def _sub(self, comp, h, has_heat):
            if self._k == self._last:
                self.g.synchronize()
                U = self.g.get_conserved()
                expected["Q"] = orc.fix_primitives(hp, orc.q_from_u(hp, U), U)
                expected["comp"] = comp
            super()._sub(comp, h, has_heat)
            if self._k == self._last and comp == "rad":
                expected["Q"][2] = self.rad.state()[1][0]
            self._k += 1

        def step(self):
            self._k, self._last = 0, 2 * len(self.comps) - 2
            return super().step()

    loop = Recorder(g1, cooling=cool1, radiation=rad1)
    g2.pause_at_checkpoints(True)
    comps_last = []
    for step in range(1, nsteps + 1):
        dt1 = loop.step()
        rad2.run_full(1, cooling=cool2)
        t2, dt2, n2 = g2.clock()
        assert n2 == step and dt2 == dt1, (step, dt1, dt2)
        g1.synchronize()
        np.testing.assert_array_equal(g2.get_conserved(), g1.get_conserved(), err_msg=f"step {step}")
        R1, h1 = rad1.state()
        R2, h2 = rad2.state()
        np.testing.assert_array_equal(R2, R1)
        np.testing.assert_array_equal(h2, h1)
        Qp = g2.print_state()
        np.testing.assert_allclose(Qp, expected["Q"], rtol=1e-13, atol=0, err_msg=f"step {step}")
        comps_last.append(expected["comp"])
    assert "rad" in comps_last


def test_star_device_loop_pauses_at_checkpoints(gpu, tmp_path):
    """Checkpoints at every step (tmax = 10 first steps, 10 checkpoints):
    the device loop pauses before each step, prints, and resumes; a paused
    step keeps calculateTimeStep's first-call rule for the step that runs.
    The state and the printed files match the host-driven loop's."""
    from torch_amd import fullstep

    _, g0, _, _, _, p0 = _star_run(gpu)
    tmax = 10 * p0.dt_max * 1.0e-20
    _, g1, cool1, rad1, _, _ = _star_run(gpu, tmax=tmax, dt_max=p0.dt_max)
    _, g2, cool2, rad2, _, _ = _star_run(gpu, tmax=tmax, dt_max=p0.dt_max)
    loop = fullstep.FullStep(g1, cooling=cool1, radiation=rad1)
    nsteps = 0
    while loop.time < tmax and nsteps < 20:
        loop.step()
        nsteps += 1
    files = gpu.run_printing(g2, 20, str(tmp_path), SCALES, cooling=cool2, radiation=rad2)
    t2, _, n2 = g2.clock()
    assert n2 == nsteps and t2 == loop.time
    assert len(files) == 11
    g1.synchronize()
    np.testing.assert_array_equal(g2.get_conserved(), g1.get_conserved())
    np.testing.assert_array_equal(rad2.state()[0], rad1.state()[0])
