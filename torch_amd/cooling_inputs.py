"""Synthetic inputs for the cooling source step on independent cells.

The workload: 2^24 independent cells with density log-uniform in
1e-24..1e-19 g/cm^3, temperature log-uniform in 10..1e5 K, HII fraction
uniform in [0, 1], ionising photon flux log-uniform in 1e6..1e14 cm^-2 s^-1,
optical depth tau uniform in [0, 10] and dt log-uniform in 1e7..1e11 s.

Mapping onto the inputs of Thermodynamics::preTimeStepCalculations /
integrate (reference src/Integrators/Thermodynamics.cpp:218-337):
  p       = rho * R * mu_inv * T, mu_inv = X (x + 1) + (1 - X)/4   (Fluid.cpp:71-74)
  E       = p / (gamma - 1)           (v = 0, so U[PRE] is the thermal energy)
  F_FUV   = 0.5 * F_ion / 1.2e7       (fluxFUV(0.5 S, r^2) with F_ion = S / (4 pi r^2))
  col_den = tau / sigma_dust          (so that Av = 1.086 sigma_dust col_den = 1.086 tau)
  ADV     = 1                         (the thermo_hii_switch gate passes)
  T_min   = 100 K                     (min_temp_initial_state = false)
All values are returned in code units for the given Converter-like object
(anything with ``to_code(v, m, l, t)``).
"""
from __future__ import annotations

import numpy as np

# Scales of config/torch-config.lua:8-10.
DENSITY_SCALE = 1.0e-20
PRESSURE_SCALE = 1.0e-28
TIME_SCALE = 1.0e11
R_CGS = 8.314462e7
SIGMA_DUST_CGS = 5.0e-22


def sample_cgs(n: int, seed: int = 2024) -> dict:
    rng = np.random.Generator(np.random.PCG64(seed))
    lu = lambda lo, hi: np.exp(rng.uniform(np.log(lo), np.log(hi), n))
    return {
        "rho": lu(1e-24, 1e-19),
        "T": lu(10.0, 1e5),
        "x": rng.uniform(0.0, 1.0, n),
        "F_ion": lu(1e6, 1e14),
        "tau": rng.uniform(0.0, 10.0, n),
        "dt": lu(1e7, 1e11),
    }


def sample_cgs_cloud(n: int, seed: int = 3030) -> dict:
    """Cells of the operator-split cloud's conditions (the 0.5 pc box of
    torch_amd.cloud): hydrogen densities from the cloud's rim to past its
    core (log-uniform 1e-22..3e-16 g/cm^3), T log-uniform 10..1e5 K, x
    exactly 0 (neutral) for a quarter, exactly 1 (the ionised sphere) for a
    quarter and uniform in [0, 1] for the rest, an unlit half (F_ion = 0: the
    ionising flux is absorbed within a cell of the source) beside a lit half
    log-uniform 1e6..1e14 cm^-2 s^-1, optical depths from none to behind the
    core (tau = 0 for 5 %, uniform in [0, 10] for 45 %, log-uniform
    10..3e3 for the rest), and dt log-uniform 1e8..1e13 s (the loop's dt is
    ~2e9 s; 1e13 s is ~5000 times it).  The exact end values put the
    cloud's own x = 0, x = 1, F = 0 and tau = 0 cells inside the training
    envelope."""
    rng = np.random.Generator(np.random.PCG64(seed))
    lu = lambda lo, hi, m: np.exp(rng.uniform(np.log(lo), np.log(hi), m))  # noqa: E731
    ux = rng.uniform(0.0, 1.0, n)
    x = np.where(ux < 0.25, 0.0, np.where(ux < 0.5, 1.0, rng.uniform(0.0, 1.0, n)))
    lit = rng.uniform(0.0, 1.0, n) < 0.5
    ut = rng.uniform(0.0, 1.0, n)
    tau = np.where(ut < 0.05, 0.0, np.where(ut < 0.5, rng.uniform(0.0, 10.0, n), lu(10.0, 3e3, n)))
    return {
        "rho": lu(1e-22, 3e-16, n),
        "T": lu(10.0, 1e5, n),
        "x": x,
        "F_ion": np.where(lit, lu(1e6, 1e14, n), 0.0),
        "tau": tau,
        "dt": lu(1e8, 1e13, n),
    }


SAMPLERS = {"c3": sample_cgs, "cloud": sample_cgs_cloud}


def to_code_inputs(s: dict, to_code, gamma: float = 1.67, X: float = 1.0, T_min: float = 100.0) -> np.ndarray:
    """(n, 9) array: den, pre, hii, adv, E, T_min, col_den, F_FUV, dt in code
    units, the column order of the cooling kernels and of the oracle."""
    n = s["rho"].shape[0]
    R = to_code(R_CGS, 0, 2, -2)
    sig = to_code(SIGMA_DUST_CGS, 0, 2, 0)
    den = s["rho"] * to_code(1.0, 1, -3, 0)
    mu_inv = X * (s["x"] + 1.0) + (1.0 - X) * 0.25
    pre = den * R * mu_inv * s["T"]
    out = np.empty((n, 9))
    out[:, 0] = den
    out[:, 1] = pre
    out[:, 2] = s["x"]
    out[:, 3] = 1.0
    out[:, 4] = pre / (gamma - 1.0)
    out[:, 5] = T_min
    out[:, 6] = s["tau"] / sig
    out[:, 7] = 0.5 * (s["F_ion"] * to_code(1.0, 0, -2, -1)) / 1.2e7
    out[:, 8] = s["dt"] * to_code(1.0, 0, 0, 1)
    return out
