"""ctypes binding of the CPU FP64 oracle (oracle/torch_oracle.c).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py`` as the checker; the product package
``torch_amd`` never imports it.

The oracle restates the reference's FP64 CPU path (hgsteggles/TORCH
``src/Integrators/{Hydro,Riemann,SlopeLimiter,Thermodynamics,SplineData}.cpp``,
``src/Fluid/{Fluid,GridCell,Grid}.cpp``, ``src/Torch/{Torch,Converter,
Constants,Parameters}.cpp``) operation by operation; the citations are in
``torch_oracle.c``.  How each function is pinned is listed in
``torch_oracle.h``.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_HARNESS = os.path.join(HERE, "_ref", "ref_harness")

DEN, PRE, HII, ADV, VEL = 0, 1, 2, 3, 4
BC = {"free": 0, "reflecting": 1, "outflow": 2, "inflow": 3, "periodic": 4}
SOLVER = {"HLL": 0, "HLLC": 1, "RotatedHLLC": 2, "default": 2}
LIMITER = {"monotonised_central": 0, "superbee": 1, "minmod": 2, "maxmod": 3,
           "leer": 4, "ospre": 5, "albada": 6, "default": 6}


class HydroParams(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
        ("spatial_order", ctypes.c_int), ("solver", ctypes.c_int), ("limiter", ctypes.c_int),
        ("bc_left", ctypes.c_int * 3), ("bc_right", ctypes.c_int * 3),
        ("dx", ctypes.c_double), ("gamma", ctypes.c_double),
        ("dfloor", ctypes.c_double), ("pfloor", ctypes.c_double), ("tfloor", ctypes.c_double),
        ("mass_frac_h", ctypes.c_double), ("spec_gas_const", ctypes.c_double),
        ("dt_max", ctypes.c_double), ("nthreads", ctypes.c_int),
    ]


class Converter(ctypes.Structure):
    _fields_ = [("M", ctypes.c_double), ("L", ctypes.c_double), ("T", ctypes.c_double)]


class ThermoParams(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in (
        "hydrogen_mass", "spec_gas_const", "boltzmann", "dust_xsec", "pi",
        "z0", "T1", "T2", "T3", "T4", "imlc", "nmlc", "ciec_minT", "ciec", "cxhi_damp", "n0", "nmc",
        "fuvh_a", "fuvh_b", "fuvh_c", "irh_a", "irh_b", "crh", "T_min_default")] + [
        ("spl_x", ctypes.c_double * 26), ("spl_y", ctypes.c_double * 26), ("spl_y2", ctypes.c_double * 26),
        ("spl_min_slope", ctypes.c_double), ("spl_max_slope", ctypes.c_double),
        ("thermo_hii_switch", ctypes.c_double), ("heating_amplification", ctypes.c_double),
        ("mass_frac_h", ctypes.c_double), ("gamma", ctypes.c_double), ("pfloor", ctypes.c_double),
        ("subcycling", ctypes.c_int)]


_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.POINTER
        d = ctypes.c_double
        dp = P(d)
        L.orc_limiter.restype = d
        L.orc_limiter.argtypes = [ctypes.c_int, d, d]
        L.orc_riemann_batch.argtypes = [ctypes.c_int, ctypes.c_int, dp, dp, P(ctypes.c_int), d, dp, P(ctypes.c_int)]
        L.orc_globalQfromU.argtypes = [P(HydroParams), dp, dp]
        L.orc_fixPrimitives.argtypes = [P(HydroParams), dp, dp]
        L.orc_fixSolution.argtypes = [P(HydroParams), dp]
        L.orc_fixSolution.restype = ctypes.c_int
        L.orc_hydro_timestep.argtypes = [P(HydroParams), dp]
        L.orc_hydro_timestep.restype = d
        L.orc_hydro_rhs.argtypes = [P(HydroParams), dp, dp]
        L.orc_hydro_rhs.restype = ctypes.c_int
        L.orc_hydro_step.argtypes = [P(HydroParams), dp, dp, dp, dp, d]
        L.orc_hydro_step.restype = ctypes.c_int
        ip = ctypes.POINTER(ctypes.c_int)
        L.orc_run_hydro.argtypes = [P(HydroParams), dp, ctypes.c_int, d, ctypes.c_int, dp, dp, ctypes.c_int, ip, ip]
        L.orc_run_hydro.restype = ctypes.c_int
        L.orc_run_hydro_print.argtypes = [P(HydroParams), dp, ctypes.c_int, d, d, ctypes.c_int, dp, dp, dp,
                                          ctypes.c_int, ip, dp, ip, ctypes.c_int, ip, ip]
        L.orc_run_hydro_print.restype = ctypes.c_int
        L.orc_UfromQ.argtypes = [dp, dp, d]
        L.orc_QfromU.argtypes = [dp, dp, d]
        L.orc_FfromQ.argtypes = [dp, dp, d, ctypes.c_int]
        L.orc_converter_set_rho_pressure_time.argtypes = [P(Converter), d, d, d]
        L.orc_to_code_units.argtypes = [P(Converter), d, d, d, d]
        L.orc_to_code_units.restype = d
        L.orc_from_code_units.argtypes = [P(Converter), d, d, d, d]
        L.orc_from_code_units.restype = d
        L.orc_thermo_init.argtypes = [P(ThermoParams), P(Converter), d, d, d, d, d, ctypes.c_int]
        L.orc_log_spline.argtypes = [P(ThermoParams), d]
        L.orc_log_spline.restype = d
        L.orc_thermo_rates.argtypes = [P(ThermoParams), d, d, d, d, d, dp]
        L.orc_thermo_batch.argtypes = [P(ThermoParams), ctypes.c_int64, dp, dp, ctypes.c_int]
        i3 = ctypes.POINTER(ctypes.c_int)
        i64 = ctypes.POINTER(ctypes.c_int64)
        L.orc_star_geometry.argtypes = [i3, i3, d, d, i64, dp, dp, dp]
        L.orc_star_wind.argtypes = [i3, i3, ctypes.c_int, ctypes.POINTER(ctypes.c_ubyte)]
        L.orc_thermo_ray_trace.argtypes = [i3, i3, i64, dp, dp, d, dp, dp, dp]
        L.orc_star_fuv.argtypes = [i3, i3, d, d, d, dp]
        L.orc_fill_heating.argtypes = [ctypes.POINTER(ThermoParams), ctypes.c_int64, dp, dp, dp,
                                       ctypes.POINTER(ctypes.c_ubyte), dp, dp]
        L.orc_fill_heating.restype = None
        L.orc_rad_init.argtypes = [ctypes.POINTER(RadParams), ctypes.POINTER(Converter), dp, d, d, d, d, d, d,
                                   ctypes.c_int, ctypes.c_int, ctypes.c_int, d, d, d, d]
        L.orc_rad_recombination.argtypes = [ctypes.POINTER(RadParams), d]
        L.orc_rad_recombination.restype = d
        L.orc_first_nonwind_cell.argtypes = [i3, i3, ctypes.c_int]
        L.orc_first_nonwind_cell.restype = ctypes.c_int64
        L.orc_rad_init_field.argtypes = [ctypes.POINTER(RadParams), i3, i3, d, d, dp, dp]
        L.orc_rad_run.argtypes = [ctypes.POINTER(RadParams), P(HydroParams), i3, ctypes.c_int, d,
                                  ctypes.POINTER(ctypes.c_int), dp, dp, dp, ctypes.c_int, dp]
        L.orc_rad_run.restype = ctypes.c_int
        _lib = L
    return _lib


def _dp(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a: np.ndarray):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def make_params(shape, dx, gamma=1.67, dfloor=1e-30, pfloor=0.0, tfloor=0.0, mass_frac_h=1.0,
                spec_gas_const=8.314462e7, dt_max=100.0, solver="RotatedHLLC", limiter="albada",
                spatial_order=1, bc="outflow", nthreads=1) -> HydroParams:
    """shape = (nz, ny, nx) of the SoA arrays."""
    p = HydroParams()
    p.nz, p.ny, p.nx = shape
    p.spatial_order = spatial_order
    p.solver = SOLVER[solver]
    p.limiter = LIMITER[limiter]
    bcs = bc if isinstance(bc, (list, tuple)) else [bc] * 6
    for i in range(3):
        p.bc_left[i] = BC[bcs[2 * i]]
        p.bc_right[i] = BC[bcs[2 * i + 1]]
    p.dx, p.gamma, p.dfloor, p.pfloor, p.tfloor = dx, gamma, dfloor, pfloor, tfloor
    p.mass_frac_h, p.spec_gas_const, p.dt_max = mass_frac_h, spec_gas_const, dt_max
    p.nthreads = nthreads
    return p


def UfromQ_grid(Q: np.ndarray, gamma: float) -> np.ndarray:
    """Fluid::globalUfromQ on a (7, ...) array (cell by cell)."""
    L = lib()
    flat = np.ascontiguousarray(Q.reshape(7, -1).T)
    out = np.empty_like(flat)
    for i in range(flat.shape[0]):
        L.orc_UfromQ(_dp(out[i]), _dp(flat[i]), gamma)
    return np.ascontiguousarray(out.T).reshape(Q.shape)


def fix_primitives(p: HydroParams, Q: np.ndarray, U: np.ndarray) -> np.ndarray:
    Q = np.ascontiguousarray(Q, dtype=np.float64).copy()
    lib().orc_fixPrimitives(ctypes.byref(p), _dp(Q), _dp(np.ascontiguousarray(U)))
    return Q


def q_from_u(p: HydroParams, U: np.ndarray) -> np.ndarray:
    U = np.ascontiguousarray(U, dtype=np.float64)
    Q = np.empty_like(U)
    lib().orc_globalQfromU(ctypes.byref(p), _dp(U), _dp(Q))
    return Q


def hydro_timestep(p: HydroParams, Q: np.ndarray) -> float:
    return lib().orc_hydro_timestep(ctypes.byref(p), _dp(np.ascontiguousarray(Q)))


def hydro_rhs(p: HydroParams, Q: np.ndarray) -> np.ndarray:
    Q = np.ascontiguousarray(Q, dtype=np.float64)
    out = np.empty_like(Q)
    bad = lib().orc_hydro_rhs(ctypes.byref(p), _dp(Q), _dp(out))
    if bad:
        raise RuntimeError("RiemannSolver::solve: HII flux is NaN")
    return out


def hydro_step(p: HydroParams, U: np.ndarray, dt: float) -> np.ndarray:
    """Torch::hydroStep(dt, true) from U; returns the new U."""
    U = np.ascontiguousarray(U, dtype=np.float64).copy()
    Q = fix_primitives(p, q_from_u(p, U), U)
    W = np.empty_like(U)
    UD = np.zeros_like(U)
    err = lib().orc_hydro_step(ctypes.byref(p), _dp(U), _dp(Q), _dp(W), _dp(UD), dt)
    if err:
        raise RuntimeError(f"hydro step failed (code {err})")
    return U


def run_hydro(p: HydroParams, U: np.ndarray, nsteps: int, tmax: float = 1e3, ncheckpoints: int = 1,
              debug: bool = True, info: bool = False):
    """Torch::run for a hydro-only problem, capped at nsteps.  Returns (U,
    dts, time); dts holds zeros after the last step run.  `debug` is
    Integration.debug (default true, Parameters.hpp:50): the loop ends after
    a step whose hydro dt is at most 1e-8 tmax (Torch.cpp:321-332).  With
    `info`, returns (U, dts of the steps run, time, quitting)."""
    U = np.ascontiguousarray(U, dtype=np.float64).copy()
    dts = np.zeros(nsteps)
    t = ctypes.c_double(0)
    ns, quit_ = ctypes.c_int(0), ctypes.c_int(0)
    err = lib().orc_run_hydro(ctypes.byref(p), _dp(U), nsteps, tmax, ncheckpoints, _dp(dts), ctypes.byref(t),
                              1 if debug else 0, ctypes.byref(ns), ctypes.byref(quit_))
    if err:
        raise RuntimeError(f"hydro run failed (code {err})")
    if info:
        return U, dts[:ns.value], t.value, bool(quit_.value)
    return U, dts, t.value


def run_hydro_print(p: HydroParams, U: np.ndarray, nsteps: int, tmax: float, ncheckpoints: int, t0: float = 0.0,
                    debug: bool = True):
    """run_hydro from time t0 with the snapshots Torch::run prints
    (Torch.cpp:240-302): returns (U, dts, time, [(count, time, Q), ...])."""
    U = np.ascontiguousarray(U, dtype=np.float64).copy()
    dts = np.zeros(nsteps)
    t = ctypes.c_double(0)
    maxp = ncheckpoints + 2
    qp = np.zeros((maxp,) + U.shape)
    pc = (ctypes.c_int * maxp)()
    pt = np.zeros(maxp)
    npr = ctypes.c_int(0)
    err = lib().orc_run_hydro_print(ctypes.byref(p), _dp(U), nsteps, t0, tmax, ncheckpoints, _dp(dts), ctypes.byref(t),
                                    _dp(qp), maxp, pc, _dp(pt), ctypes.byref(npr), 1 if debug else 0, None, None)
    if err:
        raise RuntimeError(f"hydro run failed (code {err})")
    prints = [(pc[i], pt[i], qp[i]) for i in range(npr.value)]
    return U, dts, t.value, prints


def riemann(solver: str, Ql: np.ndarray, Qr: np.ndarray, dims: np.ndarray, gamma: float):
    Ql = np.ascontiguousarray(Ql, dtype=np.float64)
    Qr = np.ascontiguousarray(Qr, dtype=np.float64)
    dims = np.ascontiguousarray(dims, dtype=np.int32)
    n = Ql.shape[0]
    F = np.zeros((n, 7))
    st = np.zeros(n, dtype=np.int32)
    lib().orc_riemann_batch(SOLVER[solver], n, _dp(Ql), _dp(Qr), _ip(dims), gamma, _dp(F), _ip(st))
    return F, st


def limiter(name: str, a: float, b: float) -> float:
    return lib().orc_limiter(LIMITER[name], a, b)


def converter(dscale: float, pscale: float, tscale: float) -> Converter:
    c = Converter()
    lib().orc_converter_set_rho_pressure_time(ctypes.byref(c), dscale, pscale, tscale)
    return c


def to_code(c: Converter, v: float, mi: float, li: float, ti: float) -> float:
    return lib().orc_to_code_units(ctypes.byref(c), v, mi, li, ti)


def from_code(c: Converter, v: float, mi: float, li: float, ti: float) -> float:
    return lib().orc_from_code_units(ctypes.byref(c), v, mi, li, ti)


def thermo_params(c: Converter, thermo_hii_switch=1e-2, heating_amplification=1.0, mass_frac_h=1.0,
                  gamma=1.67, pfloor=0.0, subcycling=True) -> ThermoParams:
    t = ThermoParams()
    lib().orc_thermo_init(ctypes.byref(t), ctypes.byref(c), thermo_hii_switch, heating_amplification,
                          mass_frac_h, gamma, pfloor, int(subcycling))
    return t


def thermo_batch(t: ThermoParams, inputs: np.ndarray, nthreads: int = 1) -> np.ndarray:
    """inputs (n, 9): den, pre, hii, adv, E, T_min, col_den, F_FUV, dt (code units).
    Returns (n, 5): heat, rate_pre, rate, nsub, p_sub."""
    inputs = np.ascontiguousarray(inputs, dtype=np.float64)
    n = inputs.shape[0]
    out = np.zeros((n, 5))
    lib().orc_thermo_batch(ctypes.byref(t), n, _dp(inputs), _dp(out), nthreads)
    return out


def thermo_rates(t: ThermoParams, nH, x, T, Av, F) -> np.ndarray:
    out = np.zeros(9)
    lib().orc_thermo_rates(ctypes.byref(t), nH, x, T, Av, F, _dp(out))
    return out


def log_spline(t: ThermoParams, x: float) -> float:
    return lib().orc_log_spline(ctypes.byref(t), x)


def _i3(v):
    return (ctypes.c_int * 3)(*[int(x) for x in v])


def star_geometry(shape, star, dx: float, pi: float):
    """Grid::calculateNearestNeighbours + Radiation::cellPathLength /
    shellVolume for a star at cell `star` = (i, j, k) of a grid of `shape` =
    (nz, ny, nx): (nb int64 [nz, ny, nx, 4], w [.., 4], ds, shell_vol)."""
    nz, ny, nx = shape
    nb = np.zeros((nz, ny, nx, 4), dtype=np.int64)
    w = np.zeros((nz, ny, nx, 4))
    ds = np.zeros((nz, ny, nx))
    sv = np.zeros((nz, ny, nx))
    lib().orc_star_geometry(_i3((nx, ny, nz)), _i3(star), dx, pi, nb.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                            _dp(w), _dp(ds), _dp(sv))
    return nb, w, ds, sv


def star_wind(shape, star, wind_radius: int) -> np.ndarray:
    """CausalWind membership (bool [nz, ny, nx]) of Fluid::initialiseGrid."""
    nz, ny, nx = shape
    out = np.zeros((nz, ny, nx), dtype=np.uint8)
    lib().orc_star_wind(_i3((nx, ny, nz)), _i3(star), int(wind_radius),
                        out.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte)))
    return out.astype(bool)


def thermo_ray_trace(shape, star, geom, hydrogen_mass: float, den: np.ndarray):
    """Thermodynamics::rayTrace: (COL_DEN, DCOL_DEN) [nz, ny, nx] of the
    density field `den` (code units) with the geometry of star_geometry."""
    nz, ny, nx = shape
    nb, w, ds, _ = geom
    den = np.ascontiguousarray(den, dtype=np.float64)
    col = np.zeros((nz, ny, nx))
    dcol = np.zeros((nz, ny, nx))
    lib().orc_thermo_ray_trace(_i3((nx, ny, nz)), _i3(star), nb.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                               _dp(w), _dp(ds), hydrogen_mass, _dp(den), _dp(col), _dp(dcol))
    return col, dcol


def fill_heating(t: ThermoParams, Q: np.ndarray, heat: np.ndarray, col_den=None, F=None, wind=None) -> np.ndarray:
    """Thermodynamics::fillHeatingArrays (Thermodynamics.cpp:419-461): H
    [11][...] of Fluid::Q (Q [7][...]); heat [3][...] (TOT, RHII, EUVH kept in
    GridCell::H between calls) is updated in place."""
    shape = Q.shape[1:]
    n = int(np.prod(shape))
    Q = np.ascontiguousarray(Q, dtype=np.float64)
    assert heat.dtype == np.float64 and heat.flags.c_contiguous and heat.shape == (3,) + shape
    col = None if col_den is None else np.ascontiguousarray(col_den, dtype=np.float64)
    f = None if F is None else np.ascontiguousarray(F, dtype=np.float64)
    w = None if wind is None else np.ascontiguousarray(wind, dtype=np.uint8)
    H = np.zeros((11,) + shape)
    lib().orc_fill_heating(ctypes.byref(t), n, _dp(Q), None if col is None else _dp(col), None if f is None else _dp(f),
                           None if w is None else w.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte)), _dp(heat), _dp(H))
    return H


def star_fuv(shape, star, dx: float, pi: float, photon_rate: float) -> np.ndarray:
    """F_FUV [nz, ny, nx] of a star-on thermo preTimeStep (code units)."""
    nz, ny, nx = shape
    F = np.zeros((nz, ny, nx))
    lib().orc_star_fuv(_i3((nx, ny, nz)), _i3(star), dx, pi, photon_rate, _dp(F))
    return F


class RadParams(ctypes.Structure):
    _fields_ = ([(k, ctypes.c_double * 4) for k in ("K",)]
                + [(k, ctypes.c_double) for k in ("sigma", "alphaB", "tau0", "minX", "THI", "THII")]
                + [(k, ctypes.c_int) for k in ("scheme", "coupling", "collisions_on")]
                + [(k, ctypes.c_double) for k in ("heating_amplification", "mass_frac_h", "hydrogen_mass",
                                                  "spec_gas_const", "boltzmann", "rydberg", "voronov_A",
                                                  "voronov_X", "voronov_K", "voronov_P", "voronov_DE",
                                                  "photon_energy", "photon_rate")]
                + [("rec_x", ctypes.c_double * 31), ("rec_y", ctypes.c_double * 31), ("rec_y2", ctypes.c_double * 31),
                   ("rec_slope", ctypes.c_double * 2), ("cool_x", ctypes.c_double * 31),
                   ("cool_y", ctypes.c_double * 31), ("cool_y2", ctypes.c_double * 31),
                   ("cool_slope", ctypes.c_double * 2)])


SCHEME = {"implicit": 0, "implicit2": 1, "explicit": 2}
COUPLING = {"tti": 0, "neq": 1, "off": 2}


def rad_params(c: Converter, K=(0.0, 0.0, 0.0, 0.0), sigma=6.3e-18, alphaB=2.59e-13, tau0=0.6, minX=0.0, THI=300.0,
               THII=1e4, scheme="implicit", coupling="off", collisions_on=False, heating_amplification=1.0,
               mass_frac_h=1.0, photon_energy=0.0, photon_rate=0.0) -> RadParams:
    """Radiation::initialise with the Lua-level (cgs) values; unknown scheme /
    coupling strings fall back to implicit / off as the reference's."""
    r = RadParams()
    Ka = (ctypes.c_double * 4)(*K)
    lib().orc_rad_init(ctypes.byref(r), ctypes.byref(c), Ka, sigma, alphaB, tau0, minX, THI, THII,
                       SCHEME.get(scheme, 0), COUPLING.get(coupling, 2), int(collisions_on), heating_amplification,
                       mass_frac_h, photon_energy, photon_rate)
    return r


def first_nonwind_cell(shape, star, wind_radius: int) -> int:
    nz, ny, nx = shape
    return lib().orc_first_nonwind_cell(_i3((nx, ny, nz)), _i3(star), int(wind_radius))


def rad_init_field(r: RadParams, shape, star, dx: float, Q: np.ndarray) -> np.ndarray:
    """Radiation::initField: R [6][nz][ny][nx] (RID order) from the fixed
    primitives Q."""
    nz, ny, nx = shape
    Q = np.ascontiguousarray(Q, dtype=np.float64)
    R = np.zeros((6, nz, ny, nx))
    lib().orc_rad_init_field(ctypes.byref(r), _i3((nx, ny, nz)), _i3(star), dx, 3.14159265359, _dp(Q), _dp(R))
    return R


def rad_run(r: RadParams, p: HydroParams, star, wind_radius: int, U: np.ndarray, R: np.ndarray, nsteps: int,
            T_min: float = 100.0, first: bool = True):
    """The radiation-only step sequence (preTimeStep, dt, two Torch::subStep
    calls per step).  Returns (U, R, Q[HII, ADV] [2][...], dts, first)."""
    U = np.ascontiguousarray(U, dtype=np.float64).copy()
    R = np.ascontiguousarray(R, dtype=np.float64).copy()
    Qo = np.zeros((2,) + U.shape[1:])
    dts = np.zeros(nsteps)
    f = ctypes.c_int(int(first))
    err = lib().orc_rad_run(ctypes.byref(r), ctypes.byref(p), _i3(star), int(wind_radius), T_min, ctypes.byref(f),
                            _dp(U), _dp(R), _dp(Qo), nsteps, _dp(dts))
    if err:
        raise RuntimeError(f"radiation run failed (code {err})")
    return U, R, Qo, dts, bool(f.value)
