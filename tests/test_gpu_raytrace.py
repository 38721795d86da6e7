"""The star-on thermo preTimeStep on the GPU: the star geometry and the
causal column-density ray trace (shell wavefront, raytrace.hip) against the
reference's own values (tests/golden/raytrace.npz, written by its Grid,
Radiation and Thermodynamics units) and, on larger grids, against the
oracle's restatement (oracle/torch_oracle.c, pinned to the same fixture):
bitwise.  Also the gathered thermo inputs (F_FUV of the star, CausalWind
cells skipped) and the device operator-split loop with a star."""
import os

import numpy as np
import pytest

import oracle as orc

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "raytrace.npz")
SCALES = (1e-20, 1e-28, 1e11)


def _setup(gpu, shape, Q, star, wind_r, S):
    nz, ny, nx = shape
    c = orc.converter(*SCALES)
    side = orc.to_code(c, 3.0857e18, 0, 1, 0)
    dx = side / nx  # Grid::initialise (Grid.cpp:471-472)
    R = orc.to_code(c, 8.314462e7, 0, 2, -2)
    dfloor = orc.to_code(c, 1e-30, 1, -3, 0)
    p = gpu.make_params(shape, side_length=side, gamma=1.67, dfloor=dfloor, pfloor=0.1 * R * dfloor * 0.1,
                        tfloor=0.1, spec_gas_const=R, tmax=orc.to_code(c, 1e13, 0, 0, 1))
    g = gpu.Grid(p)
    g.set_primitive(np.ascontiguousarray(Q))
    cool = gpu.Cooling(gpu.thermo_params(scales=SCALES))
    cool.set_star(g, gpu.make_star(star, wind_r, S))
    return g, cool, dx


@pytest.mark.parametrize("name", ["a", "b", "c", "d"])
def test_ray_trace_matches_reference_fixture(gpu, name):
    f = np.load(GOLD)
    shape = tuple(int(x) for x in f[f"{name}_shape"])
    star = tuple(int(x) for x in f[f"{name}_star"])
    wind_r = int(f[f"{name}_wind_radius"])
    S, mh, pi, dx_ref = (float(x) for x in f[f"{name}_info"])
    o = f[f"{name}_out"].reshape(-1, 15)
    g, cool, dx = _setup(gpu, shape, f[f"{name}_Q"], star, wind_r, S)
    assert dx == dx_ref and cool.params.pi == pi and cool.params.hydrogen_mass == mh
    nb, w, ds, sv = cool.star_geometry()
    np.testing.assert_array_equal(nb, o[:, 0:4].astype(np.int64))
    np.testing.assert_array_equal(w, o[:, 4:8])
    np.testing.assert_array_equal(ds, o[:, 8])
    np.testing.assert_array_equal(sv, o[:, 9])
    in9, dcol = cool.grid_inputs(g)
    np.testing.assert_array_equal(in9[6], o[:, 10])
    np.testing.assert_array_equal(dcol, o[:, 11])
    np.testing.assert_array_equal(in9[7], orc.star_fuv(shape, star, dx, pi, S).ravel())
    wind = o[:, 14] > 0
    assert (in9[3][wind] == -1.0).all() and (in9[3][~wind] >= 0).all()


@pytest.mark.parametrize("shape,star,wind_r", [((48, 48, 48), (5, 30, 20), 2), ((17, 20, 33), (32, 19, 16), 0),
                                               ((40, 24, 31), (15, 12, 20), 3)])
def test_ray_trace_matches_oracle(gpu, shape, star, wind_r):
    from torch_amd import setups

    nz, ny, nx = shape
    n0 = max(shape)
    Qt = setups.turbulent(n0, 1.5, 1.0)[:, :nz, :ny, :nx]
    c = orc.converter(*SCALES)
    Q = np.zeros((7,) + shape)
    Q[0] = 1e-22 * Qt[0] * orc.to_code(c, 1.0, 1, -3, 0)
    Q[1] = 1e-22 * Qt[0] * 8.314462e7 * 300.0 * orc.to_code(c, 1.0, 1, -1, -2)
    Q[3] = 1.0
    S = orc.to_code(c, 4.9e48, 0, 0, -1)
    g, cool, dx = _setup(gpu, shape, Q, star, wind_r, S)
    in9, dcol = cool.grid_inputs(g)
    geom = orc.star_geometry(shape, star, dx, cool.params.pi)
    col_o, dcol_o = orc.thermo_ray_trace(shape, star, geom, cool.params.hydrogen_mass, in9[0].reshape(shape))
    np.testing.assert_array_equal(in9[6], col_o.ravel())
    np.testing.assert_array_equal(dcol, dcol_o.ravel())
    assert (col_o > 0).mean() > 0.9
    wind = orc.star_wind(shape, star, wind_r).ravel()
    assert (in9[3][wind] == -1.0).all() and (in9[3][~wind] == 1.0).all()
    # a second call (buffers now hold the previous columns) gives the same
    in9b, dcolb = cool.grid_inputs(g)
    np.testing.assert_array_equal(in9b, in9)
    np.testing.assert_array_equal(dcolb, dcol)
    # new densities set between calls (the x-face buffers hold the previous
    # call's columns), each call against the oracle
    rng = np.random.default_rng(7)
    for rep in range(3):
        Q2 = Q.copy()
        Q2[0] = Q[0] * rng.uniform(0.5, 2.0, size=shape)
        g.set_primitive(Q2)
        g.synchronize()
        in9c, dcolc = cool.grid_inputs(g)
        col_o, dcol_o = orc.thermo_ray_trace(shape, star, geom, cool.params.hydrogen_mass, in9c[0].reshape(shape))
        np.testing.assert_array_equal(in9c[6], col_o.ravel(), err_msg=f"replay {rep}")
        np.testing.assert_array_equal(dcolc, dcol_o.ravel(), err_msg=f"replay {rep}")


def test_star_outside_grid_rejected(gpu):
    g, cool, _ = _setup(gpu, (8, 8, 8), np.ones((7, 8, 8, 8)), (1, 1, 1), 0, 1.0)
    with pytest.raises(gpu.TorchGpuError) as e:
        cool.set_star(g, gpu.make_star((8, 0, 0)))
    assert e.value.code == 1


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_split_loop_with_star_device_matches_host(gpu, precision):
    """Torch::fullStep with hydro and star-on thermo (ray trace in every thermo
    preTimeStep): the device loop against the host-driven sequence, bitwise."""
    from torch_amd import cloud

    n, steps = 24, 4
    runs = []
    for mode in ("host", "device"):
        p, g, cool, th = cloud.build(n, precision)
        cool.set_star(g, gpu.make_star((n // 2, n // 2, n // 2), 1, gpu.units(cloud.SCALES, 4.9e48, 0, 0, -1)))
        if mode == "host":
            loop = cloud.HostLoop(g, cool, {}, p.dt_max, surrogate=False)
            for _ in range(steps):
                loop.step()
        else:
            cool.run_split(g, steps)
        g.synchronize()
        runs.append(g.get_conserved())
    assert np.isfinite(runs[0]).all()
    np.testing.assert_array_equal(runs[0], runs[1])
