"""GPU parity of the hydro step against the FP64 oracle (bitwise for FP64).

The oracle (oracle/torch_oracle.c) is itself pinned bitwise to the
reference's own translation units by tests/test_oracle_golden.py.
"""
import numpy as np
import pytest

from torch_amd import setups

pytestmark = pytest.mark.gpu


def _grid(gpu, shape, **kw):
    return gpu.Grid(gpu.make_params(shape, side_length=kw.pop("side_length", 1.0), **kw))


def _oparams(orc, shape, **kw):
    return orc.make_params(shape, kw.pop("dx"), nthreads=8, **kw)


def _u0(orc, p, Q):
    return orc.UfromQ_grid(orc.fix_primitives(p, Q, np.zeros_like(Q)), p.gamma)


def _run_both(gpu, orc, Q, nsteps, gkw=None, okw=None):
    gkw = gkw or {}
    okw = okw or {}
    shape = Q.shape[1:]
    g = _grid(gpu, shape, **gkw)
    g.set_primitive(np.ascontiguousarray(Q))
    p = _oparams(orc, shape, dx=gkw.get("side_length", 1.0) / shape[2], **okw)
    U0 = _u0(orc, p, Q)
    np.testing.assert_array_equal(g.get_conserved(), U0)
    g.run(nsteps)
    try:
        Uo, dts, t = orc.run_hydro(p, U0, nsteps)
    except RuntimeError as e:
        # the reference throws here; the GPU path must report the same error
        code = {"1": 3, "2": 4, "3": 5}[str(e).split("code ")[1].rstrip(")")]
        with pytest.raises(gpu.TorchGpuError) as ge:
            g.synchronize()
        assert ge.value.code == code, (ge.value, e)
        return None, None, None, None
    g.synchronize()
    Ug = g.get_conserved()
    tg, dtg, n = g.clock()
    return Ug, Uo, (tg, dtg, n), (t, dts[-1], nsteps)


@pytest.mark.parametrize("n,steps", [(16, 12), (32, 5)])
def test_sod_bitwise(gpu, orc, n, steps):
    Q = setups.sod(n)
    Ug, Uo, cg, co = _run_both(gpu, orc, Q, steps)
    np.testing.assert_array_equal(Ug, Uo)
    assert cg == co


@pytest.mark.parametrize("n,sigma,mach,steps", [(16, 1.0, 2.0, 20), (24, 1.5, 4.0, 10)])
def test_turbulent_bitwise(gpu, orc, n, sigma, mach, steps):
    Q = setups.turbulent(n, sigma, mach)
    kw = dict(tfloor=1e-12, pfloor=0.1 * 8.314462e7 * 1e-30 * 1e-12)
    Ug, Uo, cg, co = _run_both(gpu, orc, Q, steps, gkw=dict(kw), okw=dict(kw))
    np.testing.assert_array_equal(Ug, Uo)
    assert cg == co


@pytest.mark.parametrize("solver,limiter,order", [
    ("HLL", "minmod", 1), ("HLLC", "superbee", 1), ("RotatedHLLC", "monotonised_central", 1),
    ("RotatedHLLC", "leer", 1), ("HLLC", "ospre", 1), ("RotatedHLLC", "maxmod", 1), ("RotatedHLLC", "albada", 0)])
def test_solver_limiter_variants_ragged_grid(gpu, orc, solver, limiter, order):
    Q = setups.turbulent(16, 1.0, 1.5)[:, :9, :10, :12].copy()  # ragged 12 x 10 x 9
    bcs = ["outflow", "inflow", "periodic", "periodic", "reflecting", "free"]
    kw = dict(tfloor=1e-12, pfloor=0.1 * 8.314462e7 * 1e-30 * 1e-12, solver=solver, limiter=limiter,
              spatial_order=order, bc=bcs)
    Ug, Uo, cg, co = _run_both(gpu, orc, Q, 6, gkw=dict(kw), okw=dict(kw))
    np.testing.assert_array_equal(Ug, Uo)


def test_explicit_step_and_timestep(gpu, orc):
    Q = setups.turbulent(16, 1.0, 2.0)
    kw = dict(tfloor=1e-12, pfloor=0.1 * 8.314462e7 * 1e-30 * 1e-12)
    g = _grid(gpu, Q.shape[1:], **kw)
    g.set_primitive(Q)
    p = _oparams(orc, Q.shape[1:], dx=1.0 / 16, **kw)
    U = _u0(orc, p, Q)
    for _ in range(3):
        dt_o = orc.hydro_timestep(p, orc.fix_primitives(p, orc.q_from_u(p, U), U))
        dt_g = g.hydro_timestep()
        assert dt_g == dt_o
        g.hydro_step(dt_g)
        U = orc.hydro_step(p, U, dt_o)
        g.synchronize()
        np.testing.assert_array_equal(g.get_conserved(), U)
    Qg = g.get_primitive(fix=True)
    np.testing.assert_array_equal(Qg, orc.fix_primitives(p, orc.q_from_u(p, U), U))


def test_deterministic_rerun(gpu):
    Q = setups.turbulent(32, 1.0, 2.0)
    kw = dict(tfloor=1e-12, pfloor=0.1 * 8.314462e7 * 1e-30 * 1e-12)
    out = []
    for _ in range(2):
        g = _grid(gpu, Q.shape[1:], **kw)
        g.set_primitive(Q)
        g.run(8)
        g.synchronize()
        out.append(g.get_conserved())
        g.close()
    np.testing.assert_array_equal(out[0], out[1])


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_nan_state_raises(gpu, precision):
    """A non-finite conserved value makes the reference's checkValues throw
    (Torch.cpp:420-443); the GPU path reports it at the next sync, and the
    run ends at the failing step as the reference's does."""
    Q = setups.sod(16)
    g = _grid(gpu, Q.shape[1:], precision=gpu.FP64 if precision == "fp64" else gpu.MIXED)
    g.set_primitive(Q)
    U = g.get_conserved()
    U[1, 3, 4, 5] = np.nan
    g.set_conserved(U)
    g.run(5)
    with pytest.raises(gpu.TorchGpuError) as e:
        g.synchronize()
    assert e.value.code in (4, 5)
    assert "(step 0)" in str(e.value)
    assert g.clock()[2] == 1


def test_invalid_order_rejected(gpu):
    with pytest.raises(gpu.TorchGpuError) as e:
        gpu.Grid(gpu.make_params((8, 8, 8), spatial_order=2))
    assert e.value.code == 6


def test_mixed_precision_close_to_fp64(gpu):
    Q = setups.turbulent(32, 1.0, 2.0)
    kw = dict(tfloor=1e-12, pfloor=0.1 * 8.314462e7 * 1e-30 * 1e-12)
    res = []
    for prec in (gpu.FP64, gpu.MIXED):
        g = _grid(gpu, Q.shape[1:], precision=prec, **kw)
        g.set_primitive(Q)
        g.run(20)
        g.synchronize()
        res.append(g.get_primitive(fix=False))
        g.close()
    d64, dmx = res[0][0], res[1][0]
    rel_l1 = np.abs(dmx - d64).sum() / np.abs(d64).sum()
    assert rel_l1 < 1e-5, rel_l1


@pytest.mark.parametrize("solver,limiter,order", [
    ("HLL", "minmod", 1), ("HLLC", "superbee", 1), ("RotatedHLLC", "albada", 1), ("RotatedHLLC", "leer", 1),
    ("RotatedHLLC", "albada", 0)])
def test_mixed_kernel_variants_ragged_grid(gpu, orc, solver, limiter, order):
    """The mixed-precision kernel (float reconstruction and flux, rotation-free
    hybrid solver, FP64 accumulation) on a ragged 12 x 10 x 9 grid with every
    boundary condition against the FP64 oracle: relative L1 of density and
    pressure below 1e-5 after 6 steps (FP32 rounding, ~1e-7 per step)."""
    Q = setups.turbulent(16, 1.0, 1.5)[:, :9, :10, :12].copy()
    bcs = ["outflow", "inflow", "periodic", "periodic", "reflecting", "free"]
    kw = dict(tfloor=1e-12, pfloor=0.1 * 8.314462e7 * 1e-30 * 1e-12, solver=solver, limiter=limiter,
              spatial_order=order, bc=bcs)
    g = _grid(gpu, Q.shape[1:], precision=gpu.MIXED, **kw)
    g.set_primitive(np.ascontiguousarray(Q))
    p = _oparams(orc, Q.shape[1:], dx=1.0 / Q.shape[3], **kw)
    Uo, dts, t = orc.run_hydro(p, _u0(orc, p, Q), 6)
    g.run(6)
    g.synchronize()
    Ug = g.get_conserved()
    for v in (0, 1):
        rel = np.abs(Ug[v] - Uo[v]).sum() / np.abs(Uo[v]).sum()
        assert rel < 1e-5, (v, rel)
    tg, dtg, n = g.clock()
    assert n == 6 and abs(tg - t) / t < 1e-5


def test_mixed_kernel_interior_and_ring_columns(gpu, orc):
    """A 64 x 48 x 40 grid, large enough that the mixed kernel has columns
    whose x and y halos lie inside the grid (their loads skip the ghost
    mapping; 16 x 16 columns: 2 x 1 of them) beside boundary columns, with
    periodic, reflecting, outflow and inflow boundaries and 40 planes (three
    z blocks, the last one partial): relative L1 of density, energy and
    momenta against the FP64 oracle below 1e-5 after 4 steps."""
    from torch_amd import setups

    Q = setups.turbulent(64, 1.0, 1.5)[:, :40, :48, :64].copy()
    bcs = ["periodic", "periodic", "reflecting", "outflow", "inflow", "outflow"]
    kw = dict(tfloor=1e-12, pfloor=0.1 * 8.314462e7 * 1e-30 * 1e-12, bc=bcs)
    g = _grid(gpu, Q.shape[1:], precision=gpu.MIXED, **kw)
    g.set_primitive(np.ascontiguousarray(Q))
    p = _oparams(orc, Q.shape[1:], dx=1.0 / Q.shape[3], **kw)
    Uo, dts, t = orc.run_hydro(p, _u0(orc, p, Q), 4)
    g.run(4)
    g.synchronize()
    Ug = g.get_conserved()
    for v in (0, 1, 4, 6):
        rel = np.abs(Ug[v] - Uo[v]).sum() / np.abs(Uo[v]).sum()
        assert rel < 1e-5, (v, rel)


@pytest.mark.parametrize("solver,limiter", [("RotatedHLLC", "albada"), ("HLL", "ospre"), ("HLLC", "leer")])
def test_mixed_kernel_large_code_values(gpu, orc, solver, limiter):
    """Code units with pressures of 1e24 and velocities of 1e12 (the
    turbulent field with p x 1e24, v x 1e12, the gas constant x 1e24 so that
    temperatures are unchanged): the squares of the FP32 limiter and the HLL
    S_l S_r dU term would overflow FP32 there.  The mixed kernel must match
    the FP64 oracle on the scaled problem to the same 1e-5 relative L1 as on
    the unscaled one."""
    Q = setups.turbulent(16, 1.0, 1.5)[:, :9, :10, :12].copy()
    Q[1] *= 1e24
    Q[4:7] *= 1e12
    R = 8.314462e7 * 1e24
    kw = dict(tfloor=1e-12, pfloor=0.1 * R * 1e-30 * 1e-12, spec_gas_const=R, solver=solver, limiter=limiter,
              bc=["outflow", "inflow", "periodic", "periodic", "reflecting", "free"])
    g = _grid(gpu, Q.shape[1:], precision=gpu.MIXED, **kw)
    g.set_primitive(np.ascontiguousarray(Q))
    p = _oparams(orc, Q.shape[1:], dx=1.0 / Q.shape[3], **kw)
    Uo, dts, t = orc.run_hydro(p, _u0(orc, p, Q), 6)
    assert np.isfinite(Uo).all()
    g.run(6)
    g.synchronize()
    Ug = g.get_conserved()
    for v in (0, 1):
        rel = np.abs(Ug[v] - Uo[v]).sum() / np.abs(Uo[v]).sum()
        assert rel < 1e-5, (v, rel)


def test_sod64_200_steps_bitwise_vs_reference_fixture(gpu):
    """The 64^3 Sod configuration through 200 steps on the GPU FP64 path
    against tests/golden/sod64.npz, which the reference's own translation
    units wrote: one x-line at steps 1, 2, 10 and 200 (all lines of the
    planar problem must be identical) and every dt."""
    import os

    f = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "sod64.npz"))
    g = _grid(gpu, (64, 64, 64))
    g.set_primitive(np.ascontiguousarray(setups.sod(64)))
    dts = []
    for s in range(1, 201):
        g.run(1)
        _, dt, n = g.clock()
        assert n == s
        dts.append(dt)
        if s in (1, 2, 10, 200):
            U = g.get_conserved()
            np.testing.assert_array_equal(U[:, 0, 0, :], f[f"U_{s}"], err_msg=f"step {s}")
            assert (U == U[:, :1, :1, :]).all(), s
    np.testing.assert_array_equal(np.array(dts), f["dts"])


def test_mixed_kernel_64bit_addressing_path(gpu):
    """The mixed kernel addresses each state variable by a 32-bit byte offset
    below 2^29 cells and with 64-bit addresses above; TORCHGPU_ADDR64=1 (read
    when a grid is created) takes the 64-bit path on a small grid.  Only the
    addressing differs, so both give the same state bit for bit (ragged grid
    with every boundary condition, 6 steps)."""
    import os

    Q = np.ascontiguousarray(setups.turbulent(32, 1.0, 2.0)[:, :20, :27, :31])
    bcs = ["outflow", "inflow", "periodic", "periodic", "reflecting", "free"]
    kw = dict(tfloor=1e-12, pfloor=0.1 * 8.314462e7 * 1e-30 * 1e-12, bc=bcs)
    out = []
    for a64 in ("0", "1"):
        os.environ["TORCHGPU_ADDR64"] = a64
        try:
            g = _grid(gpu, Q.shape[1:], precision=gpu.MIXED, **kw)
        finally:
            os.environ.pop("TORCHGPU_ADDR64", None)
        g.set_primitive(Q)
        g.run(6)
        g.synchronize()
        out.append((g.get_conserved(), g.clock()))
        g.close()
    assert np.array_equal(out[0][0], out[1][0]) and out[0][1] == out[1][1]
