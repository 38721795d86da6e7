"""The mixed-precision hydro path (FP32 reconstruction and Riemann flux, FP64
accumulation, pressure recovery and CFL) on the 512^3 workload's field
statistics -- sigma_ln_rho = 1.5, rms Mach 4 -- at 24^3, against the
reference's own FP64 states (tests/golden/turb24_s15_m4.npz, written by the
reference's translation units): after 2 and 20 steps, with its own CFL dt
sequence, the relative L1 error of density is held to 1e-6 (the target the
512^3 run is held to after 200 steps), of pressure and velocity to 1e-5, and
every dt to 1e-5 relative.  The same run driven with the reference's dt
sequence (Torch::hydroStep(dt) per step) is held to the same bounds."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _prims(U, gamma=1.67):
    rho = U[0]
    v = U[4:7] / rho
    p = (U[1] - 0.5 * (U[4:7] * v).sum(0)) * (gamma - 1.0)
    return rho, p, v


def _errors(U, Ur):
    out = {}
    for name, a, b in zip(("density", "pressure", "velocity"), _prims(U), _prims(Ur)):
        d = np.abs(a - b)
        out[name] = (d.sum() / np.abs(b).sum(), d.max() / np.abs(b).max())
    return out


def _grid(gpu, f):
    Q = f["Q"]
    tfloor = float(f["tfloor"])
    g = gpu.Grid(gpu.make_params(Q.shape[1:], tfloor=tfloor, pfloor=0.1 * 8.314462e7 * 1e-30 * tfloor,
                                 precision=gpu.MIXED))
    g.set_primitive(np.ascontiguousarray(Q))
    return g


def _check(errs, tag):
    print(tag, {k: f"L1 {a:.2e} Linf {b:.2e}" for k, (a, b) in errs.items()})
    assert errs["density"][0] < 1e-6, (tag, errs)
    assert errs["pressure"][0] < 1e-5, (tag, errs)
    assert errs["velocity"][0] < 1e-5, (tag, errs)


def test_mixed_c4_field_own_dt(gpu):
    f = np.load(os.path.join(GOLD, "turb24_s15_m4.npz"))
    g = _grid(gpu, f)
    steps = int(f["steps"])
    dts = []
    for s in range(1, steps + 1):
        g.run(1)
        dts.append(g.clock()[1])
        if s == 2:
            g.synchronize()
            _check(_errors(g.get_conserved(), f["U2"]), "step 2")
    g.synchronize()
    np.testing.assert_allclose(np.array(dts), f["dts"], rtol=1e-5, atol=0)
    _check(_errors(g.get_conserved(), f["Uend"]), f"step {steps}")


def test_mixed_c4_field_reference_dt(gpu):
    f = np.load(os.path.join(GOLD, "turb24_s15_m4.npz"))
    g = _grid(gpu, f)
    for dt in f["dts"]:
        g.hydro_step(float(dt))
    g.synchronize()
    _check(_errors(g.get_conserved(), f["Uend"]), "reference dts")
