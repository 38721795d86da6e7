"""Pins the CPU oracle to the reference's own FP64 results.

Fixtures in tests/golden/ were written by the reference's unmodified
translation units (oracle/_ref harness, tests/golden/make_golden.py).  The
oracle must reproduce them bitwise: every hydro state, every dt, every
Riemann flux and every unit-converted constant.  Thermodynamics involves
glibc exp/log/pow on both sides and is also checked bitwise (same libm).
"""
import json
import os

import numpy as np
import pytest

import oracle

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    return np.load(os.path.join(GOLD, name), allow_pickle=False)


def _u0(p, Q):
    return oracle.UfromQ_grid(oracle.fix_primitives(p, Q, np.zeros_like(Q)), p.gamma)


@pytest.mark.parametrize("solver", ["HLL", "HLLC", "RotatedHLLC"])
def test_riemann_kats_bitwise(solver):
    f = _load("riemann.npz")
    F, st = oracle.riemann(solver, f["Ql"], f["Qr"], f["dims"], float(f["gamma"]))
    np.testing.assert_array_equal(F, f[f"F_{solver}"])
    np.testing.assert_array_equal(st, f[f"status_{solver}"].astype(np.int32))


@pytest.mark.parametrize("name", ["turb16_s1_m2", "turb24_s15_m4"])
def test_turbulent_steps_bitwise(name):
    f = _load(f"{name}.npz")
    Q = f["Q"]
    tfloor = float(f["tfloor"])
    p = oracle.make_params(Q.shape[1:], 1.0 / Q.shape[3], tfloor=tfloor, pfloor=0.1 * 8.314462e7 * 1e-30 * tfloor,
                           nthreads=4)
    U0 = _u0(p, Q)
    np.testing.assert_array_equal(U0, f["U0"])
    U2, d2, _ = oracle.run_hydro(p, U0, 2)
    np.testing.assert_array_equal(U2, f["U2"])
    steps = int(f["steps"])
    Ue, de, _ = oracle.run_hydro(p, U0, steps)
    np.testing.assert_array_equal(Ue, f["Uend"])
    np.testing.assert_array_equal(de, f["dts"])


def test_variants_bitwise():
    f = _load("variants.npz")
    Q = f["Q"]
    bcs = ["outflow", "inflow", "periodic", "periodic", "reflecting", "free"]
    for key in [k[2:] for k in f.files if k.startswith("U_")]:
        solver, rest = key.split("_", 1)
        limiter, order = rest.rsplit("_", 1)
        p = oracle.make_params(Q.shape[1:], 1.0 / Q.shape[3], tfloor=1e-12, pfloor=0.1 * 8.314462e7 * 1e-30 * 1e-12,
                               solver=solver, limiter=limiter, spatial_order=int(order), bc=bcs, nthreads=4)
        U, dts, _ = oracle.run_hydro(p, _u0(p, Q), 6)
        np.testing.assert_array_equal(U, f[f"U_{key}"], err_msg=key)
        np.testing.assert_array_equal(dts, f[f"dts_{key}"], err_msg=key)


@pytest.mark.skipif(not os.path.exists(os.path.join(GOLD, "sod64.npz")), reason="sod64 fixture not generated")
def test_sod64_200_steps_bitwise():
    """Config 1: 64^3 Sod, 200 steps.  The reference's (j, k) lines are all
    identical, so one line of the oracle is compared at steps 1, 2, 10, 200
    and the whole dt sequence."""
    f = _load("sod64.npz")
    from torch_amd import setups

    Q = setups.sod(64, 2, 2)  # planar: two transverse cells carry the same line
    p = oracle.make_params(Q.shape[1:], 1.0 / 64, bc="outflow", nthreads=4)
    U0 = _u0(p, Q)
    U = U0
    done = 0
    dts_all = []
    for s in (1, 2, 10, 200):
        # restart-free: rerun from U0 to step s (run_hydro starts a new clock)
        U, dts, t = oracle.run_hydro(p, U0, s)
        np.testing.assert_array_equal(U[:, 0, 0, :], f[f"U_{s}"], err_msg=f"step {s}")
        assert (U == U[:, :1, :1, :]).all()
        dts_all = dts
    np.testing.assert_array_equal(dts_all, f["dts"])


def test_units_and_constants_bitwise():
    with open(os.path.join(GOLD, "consts.json")) as fh:
        ref = json.load(fh)
    for name, r in ref.items():
        c = oracle.converter(*r["scales"])
        assert (c.M, c.L, c.T) == (r["M"], r["L"], r["T"]), name
        tc = lambda v, m, l, t: oracle.to_code(c, v, m, l, t)  # noqa: E731
        assert tc(1.674e-24, 1, 0, 0) == r["hydrogen_mass"]
        assert tc(8.314462e7, 0, 2, -2) == r["spec_gas_const"]
        assert tc(1.3806488e-16, 1, 2, -2) == r["boltzmann"]
        assert tc(5.0e-22, 0, 2, 0) == r["dust_xsec"]
        dfloor = tc(1e-30, 1, -3, 0)
        assert dfloor == r["dfloor"]
        assert 0.1 * r["spec_gas_const"] * dfloor * 0.1 == r["pfloor"]
        t = oracle.thermo_params(c)
        for k in ("imlc", "nmlc", "ciec", "n0", "nmc", "fuvh_a", "fuvh_b", "fuvh_c", "irh_a", "irh_b", "crh"):
            assert getattr(t, k) == r[k], (name, k)
        assert list(t.spl_x) == r["spl_x"]
        assert list(t.spl_y) == r["spl_y"]
        assert list(t.spl_y2) == r["spl_y2"]
        assert (t.spl_min_slope, t.spl_max_slope) == (r["spl_min_slope"], r["spl_max_slope"])


def test_thermo_cells_match_reference():
    f = _load("thermo.npz")
    c = oracle.converter(1e-20, 1e-28, 1e11)
    t = oracle.thermo_params(c, pfloor=float(f["pfloor"]))
    out = oracle.thermo_batch(t, f["inputs"], nthreads=4)
    np.testing.assert_array_equal(out[:, 0][f["inputs"][:, 3] >= 1e-2], f["heat"][f["inputs"][:, 3] >= 1e-2])
    np.testing.assert_array_equal(out[:, 1], f["rate_pre"])
    np.testing.assert_array_equal(out[:, 2], f["rate"])
    # gate: ADV below thermo_hii_switch -> RATE exactly 0 (Thermodynamics.cpp:227-230,275-280)
    gated = f["inputs"][:, 3] < 1e-2
    assert gated.sum() == 100 and (out[gated, 2] == 0).all()


@pytest.mark.parametrize("case", ["early", "late", "nodebug", "run"])
def test_debug_stop_bitwise(case):
    """Torch::calculateTimeStep's debug test (Torch.cpp:321-332) against the
    reference harness's loop: the step after which the loop quits, every
    dt, the time and the state, bitwise; without debug, or with the hydro dt
    above 1e-8 tmax, every step runs."""
    f = _load("debug_stop.npz")
    Q = f["Q"]
    tfloor = float(f["tfloor"])
    tmax = float(f[f"{case}_tmax"])
    p = oracle.make_params(Q.shape[1:], 1.0 / Q.shape[3], tfloor=tfloor, pfloor=0.1 * 8.314462e7 * 1e-30 * tfloor,
                           dt_max=tmax / 10.0, nthreads=4)
    U, dts, t, quit_ = oracle.run_hydro(p, _u0(p, Q), int(f["nsteps"]), tmax=tmax,
                                        debug=bool(f[f"{case}_debug"]), info=True)
    assert len(dts) == int(f[f"{case}_steps"])
    assert quit_ == bool(f[f"{case}_quitting"])
    np.testing.assert_array_equal(dts, f[f"{case}_dts"])
    assert t == float(f[f"{case}_time"])
    np.testing.assert_array_equal(U, f[f"{case}_U"])


def test_debug_stop_planted_mistake():
    """A stop that tested dt rather than 100 dt / tmax, or tested the
    clipped dt, would quit in the 'run' case as well; one that missed the
    test would run on in 'early'.  Both are told apart by the fixture: the
    early case quits after exactly two steps while 'run' (tmax 1e3) does
    not."""
    f = _load("debug_stop.npz")
    assert int(f["early_steps"]) == 2 and int(f["run_steps"]) == int(f["nsteps"])
    # the threshold sits between the late case's stopping dt and every
    # earlier CFL dt of its run
    late = f["late_dts"]
    thr = 1e-8 * float(f["late_tmax"])
    assert late[-1] <= thr and (late[1:-1] > thr).all()
