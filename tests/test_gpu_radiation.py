"""The Radiation integrator on the GPU (radiation.hip) against the
reference's own outputs (tests/golden/radiation.npz, written by its
Radiation, Grid and Fluid units; the oracle reproduces them bitwise,
tests/test_oracle_radiation.py) and against the oracle on a larger grid.
The GPU evaluates exp / pow / log with the device math library, so the
comparison is to rounding: every dt within 1e-12 relative; U, R and
Q[HII, ADV] with median relative difference below 1e-12 and maximum below
1e-6 (the implicit HII iteration stops at a 1e-3 convergence test, so a
last-bit difference can move one cell's iteration count)."""
import os
import sys

import numpy as np
import pytest

import oracle as orc

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
from make_golden import RAD_CASES  # noqa: E402

GOLD = os.path.join(HERE, "golden", "radiation.npz")
SCALES = (1e-20, 1e-28, 1e11)


def _close(a, b, what, med=1e-12, mx=1e-6):
    a, b = np.asarray(a).ravel(), np.asarray(b).ravel()
    rel = np.abs(a - b) / np.maximum(np.abs(b), 1e-300)
    rel[(a == b)] = 0.0
    assert np.isfinite(a).all(), what
    assert np.median(rel) <= med and rel.max() <= mx, (what, np.median(rel), rel.max(), int((rel > mx).sum()))


def _run_gpu(gpu, shape, Q, side_cgs, star, wind, kv, nsteps):
    c = orc.converter(*SCALES)
    tc = lambda v, m, l, t: orc.to_code(c, v, m, l, t)  # noqa: E731
    R = tc(8.314462e7, 0, 2, -2)
    dfloor = tc(1e-30, 1, -3, 0)
    p = gpu.make_params(shape, side_length=tc(side_cgs, 0, 1, 0), gamma=1.67, dfloor=dfloor,
                        pfloor=0.1 * R * dfloor * 0.1, tfloor=0.1, spec_gas_const=R, tmax=tc(1e13, 0, 0, 1))
    g = gpu.Grid(p)
    g.set_primitive(np.ascontiguousarray(Q))
    rp = gpu.rad_params(SCALES, K=(kv.get("K1", 0.0), kv.get("K2", 0.0), kv.get("K3", 0.0), kv.get("K4", 0.0)),
                        scheme=kv["integration_scheme"], coupling=kv["coupling"],
                        collisions_on=bool(kv["collisions_on"]),
                        heating_amplification=kv.get("heating_amplification", 1.0))
    S = gpu.units(SCALES, 1e49, 0, 0, -1)
    rad = gpu.Radiation(rp, g, gpu.make_star(star, wind, S))
    rad.init_field(Q)
    dts = []
    for _ in range(nsteps):
        rad.pre_step()
        dt = min(p.dt_max, rad.timestep(p.dt_max))
        dts.append(dt)
        rad.substep(dt, True)
        rad.substep(dt, False)
    g.synchronize()
    Rg, qg = rad.state()
    return g.get_conserved(), Rg, qg, np.array(dts)


@pytest.mark.parametrize("name", list(RAD_CASES))
def test_radiation_matches_reference_fixture(gpu, name):
    f = np.load(GOLD)
    cs = RAD_CASES[name]
    U, R, q, dts = _run_gpu(gpu, cs["shape"], f[f"{name}_Q"], cs["side"], cs["star"], cs["wind"], cs["kv"],
                            cs["nsteps"])
    _close(dts, f[f"{name}_dts"], "dt", med=1e-12, mx=1e-12)
    _close(U, f[f"{name}_U"], "U")
    _close(R, f[f"{name}_R"], "R")
    _close(q, f[f"{name}_q"], "Q")


@pytest.mark.parametrize("kv", [dict(K1=0.2, coupling="neq", integration_scheme="implicit", collisions_on=0),
                                dict(K3=0.2, coupling="tti", integration_scheme="explicit", collisions_on=1)])
def test_radiation_matches_oracle_32(gpu, kv):
    """A 32 x 28 x 24 grid, star off-centre with a wind radius of 2, 3 steps."""
    shape, star, wind = (24, 28, 32), (20, 9, 11), 2
    c = orc.converter(*SCALES)
    tc = lambda v, m, l, t: orc.to_code(c, v, m, l, t)  # noqa: E731
    rng = np.random.default_rng(5)
    rho = 10.0 ** rng.uniform(-22, -20, shape)
    T = 10.0 ** rng.uniform(1.5, 4.0, shape)
    hii = rng.uniform(0, 0.05, shape)
    Q = np.zeros((7,) + shape)
    Q[0] = rho * tc(1.0, 1, -3, 0)
    Q[1] = rho * 8.314462e7 * (hii + 1.0) * T * tc(1.0, 1, -1, -2)
    Q[2] = hii
    Q[3] = 1.0
    side = 1e18
    U, R, q, dts = _run_gpu(gpu, shape, Q, side, star, wind, kv, 3)
    Rg = tc(8.314462e7, 0, 2, -2)
    dfloor = tc(1e-30, 1, -3, 0)
    p = orc.make_params(shape, tc(side, 0, 1, 0) / shape[2], gamma=1.67, dfloor=dfloor,
                        pfloor=0.1 * Rg * dfloor * 0.1, tfloor=0.1, spec_gas_const=Rg,
                        dt_max=tc(1e13, 0, 0, 1) / 10.0, nthreads=4)
    r = orc.rad_params(c, K=(kv.get("K1", 0.0), 0.0, kv.get("K3", 0.0), 0.0), scheme=kv["integration_scheme"],
                       coupling=kv["coupling"], collisions_on=bool(kv["collisions_on"]), photon_energy=2.976e-11,
                       photon_rate=1e49)
    Qf = orc.fix_primitives(p, Q, np.zeros_like(Q))
    U0 = orc.UfromQ_grid(Qf, p.gamma)
    R0 = orc.rad_init_field(r, shape, star, p.dx, Qf)
    Uo, Ro, qo, dto, _ = orc.rad_run(r, p, star, wind, U0, R0, 3)
    _close(dts, dto, "dt", med=1e-12, mx=1e-12)
    _close(U, Uo, "U")
    _close(R, Ro, "R")
    _close(q, qo, "Q")
    assert not np.array_equal(U[2], U0[2])


def test_radiation_star_outside_grid_rejected(gpu):
    g = gpu.Grid(gpu.make_params((8, 8, 8)))
    with pytest.raises(gpu.TorchGpuError) as e:
        gpu.Radiation(gpu.rad_params(), g, gpu.make_star((0, 9, 0)))
    assert e.value.code == 1
