"""Writes the golden fixtures of tests/golden/ by running the reference's own
FP64 translation units (hgsteggles/TORCH src/Integrators, src/Fluid,
src/Torch, compiled unmodified into oracle/_ref/ref_harness by
``make -C oracle ref``; the harness's driver restates Torch::run /
fullStep / hydroStep / calculateTimeStep, src/Torch/Torch.cpp:240-418).

Run in the build container (where /root/reference is mounted):
    make -C oracle ref && python tests/golden/make_golden.py
The fixtures are inputs and the reference's outputs only; no reference source
is stored.  Sizes are kept small (a few MB in total).
"""
from __future__ import annotations

import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from torch_amd import cooling_inputs, setups  # noqa: E402

HARNESS = oracle.REF_HARNESS
R_CGS = 8.314462e7


def run_hydro_case(Q: np.ndarray, nsteps: int, dumps, tfloor=0.0, extra=None):
    nz, ny, nx = Q.shape[1:]
    with tempfile.TemporaryDirectory() as d:
        Q.astype(np.float64).tofile(f"{d}/q.bin")
        kv = {"nx": nx, "ny": ny, "nz": nz, "side_length": 1.0, "gamma": 1.67, "density_floor": 1e-30,
              "temperature_floor": tfloor, "simulation_time": 1000, "ncheckpoints": 1, "nsteps": nsteps,
              "input_q": f"{d}/q.bin", "dump_steps": " ".join(str(s) for s in dumps), "output_prefix": f"{d}/ref"}
        kv.update(extra or {})
        with open(f"{d}/case.txt", "w") as f:
            for k, v in kv.items():
                f.write(f"{k} {v}\n")
        r = subprocess.run([HARNESS, "hydro", f"{d}/case.txt"], capture_output=True, text=True, check=True)
        out = {s: np.fromfile(f"{d}/ref_U_{s:04d}.bin").reshape(7, nz, ny, nx) for s in dumps}
        dts = np.fromfile(f"{d}/ref_dt.bin")
        return out, dts, json.loads(r.stdout.strip().splitlines()[-1])


def sod():
    n = 64
    Q = setups.sod(n)
    out, dts, info = run_hydro_case(Q, 200, [1, 2, 10, 200])
    lines = {}
    for s, U in out.items():
        # planar symmetry: every (j, k) line is bitwise identical
        assert (U == U[:, :1, :1, :]).all(), f"Sod step {s}: lines differ"
        lines[f"U_{s}"] = U[:, 0, 0, :].copy()
    np.savez_compressed(f"{HERE}/sod64.npz", dts=dts, time=info["time"], **lines)
    print("sod64", info)


def turbulent():
    cases = {
        "turb16_s1_m2": dict(n=16, sigma=1.0, mach=2.0, steps=20),
        "turb24_s15_m4": dict(n=24, sigma=1.5, mach=4.0, steps=20),
    }
    for name, c in cases.items():
        Q = setups.turbulent(c["n"], c["sigma"], c["mach"])
        out, dts, info = run_hydro_case(Q, c["steps"], [0, 2, c["steps"]], tfloor=1e-12)
        np.savez_compressed(f"{HERE}/{name}.npz", Q=Q, U0=out[0], U2=out[2], Uend=out[c["steps"]], dts=dts,
                            steps=c["steps"], tfloor=1e-12)
        print(name, info)


def variants():
    """Ragged 12x10x9 grid, mixed boundaries, other solvers and limiters."""
    Q = setups.turbulent(16, 1.0, 1.5)[:, :9, :10, :12].copy()
    bcs = {"left_boundary_condition_x": "outflow", "right_boundary_condition_x": "inflow",
           "left_boundary_condition_y": "periodic", "right_boundary_condition_y": "periodic",
           "left_boundary_condition_z": "reflecting", "right_boundary_condition_z": "free"}
    res = {"Q": Q}
    for solver, limiter, order in [("HLL", "minmod", 1), ("HLLC", "superbee", 1),
                                   ("RotatedHLLC", "monotonised_central", 1), ("RotatedHLLC", "leer", 1),
                                   ("HLLC", "ospre", 1), ("RotatedHLLC", "albada", 0)]:
        extra = dict(bcs, riemann_solver=solver, slope_limiter=limiter, spatial_order=order)
        out, dts, info = run_hydro_case(Q, 6, [6], tfloor=1e-12, extra=extra)
        key = f"{solver}_{limiter}_{order}"
        res[f"U_{key}"] = out[6]
        res[f"dts_{key}"] = dts
        print(key, info)
    np.savez_compressed(f"{HERE}/variants.npz", **res)


def riemann():
    rng = np.random.Generator(np.random.PCG64(7))
    n = 3000
    Ql = np.zeros((n, 7))
    Qr = np.zeros((n, 7))
    for Q in (Ql, Qr):
        Q[:, 0] = np.exp(rng.normal(0, 1, n))
        Q[:, 1] = np.exp(rng.normal(0, 1, n))
        Q[:, 2] = rng.uniform(0, 1, n)
        Q[:, 3] = rng.uniform(0, 1, n)
        Q[:, 4:7] = rng.normal(0, 2, (n, 3))
    dims = rng.integers(0, 3, n)
    # branch coverage: equal velocities (pure HLLC), velocity jump along the
    # normal only (pure HLL), supersonic left/right states
    k = n // 6
    Qr[:k, 4:7] = Ql[:k, 4:7]
    for i in range(k, 2 * k):
        Qr[i, 4:7] = Ql[i, 4:7]
        Qr[i, 4 + dims[i]] += rng.normal(0, 1)
    Ql[2 * k:3 * k, 4:7] += 10.0
    Qr[2 * k:3 * k, 4:7] += 10.0
    Ql[3 * k:4 * k, 4:7] -= 10.0
    Qr[3 * k:4 * k, 4:7] -= 10.0
    Qr[4 * k:4 * k + 50, 4:7] = Ql[4 * k:4 * k + 50, 4:7] + 1e-8  # |dv| < 1e-6 threshold
    res = {"Ql": Ql, "Qr": Qr, "dims": dims.astype(np.int32), "gamma": 1.67}
    inp = np.concatenate([Ql, Qr, dims[:, None].astype(float)], axis=1)
    with tempfile.TemporaryDirectory() as d:
        inp.tofile(f"{d}/in.bin")
        for solver in ("HLL", "HLLC", "RotatedHLLC"):
            subprocess.run([HARNESS, "riemann", solver, "1.67", str(n), f"{d}/in.bin", f"{d}/out.bin"], check=True)
            out = np.fromfile(f"{d}/out.bin").reshape(n, 8)
            res[f"F_{solver}"] = out[:, :7]
            res[f"status_{solver}"] = out[:, 7]
    np.savez_compressed(f"{HERE}/riemann.npz", **res)
    print("riemann", n)


def consts():
    res = {}
    for name, scales in {"unit": (1.0, 1.0, 1.0), "cooling": (1e-20, 1e-28, 1e11),
                         "cloud": (1e-20, 1e-28, 1e11)}.items():
        r = subprocess.run([HARNESS, "consts", *map(str, scales), "0.1", "1e-30"], capture_output=True, text=True,
                           check=True)
        res[name] = json.loads(r.stdout)
        res[name]["scales"] = scales
    with open(f"{HERE}/consts.json", "w") as f:
        json.dump(res, f, indent=1)
    print("consts")


def thermo():
    """Per-cell Thermodynamics::preTimeStepCalculations + integrate on cells
    drawn from the cooling-workload distribution (code units of the
    1e-20/1e-28/1e11 scales).  Cells whose subcycle count is large are
    thinned so that the reference finishes in about a minute."""
    c = oracle.converter(1e-20, 1e-28, 1e11)
    tc = lambda v, m, l, t: oracle.to_code(c, v, m, l, t)  # noqa: E731
    s = cooling_inputs.sample_cgs(6000, seed=11)
    inp9 = cooling_inputs.to_code_inputs(s, tc)
    # estimate subcycle counts with the oracle (subcycling off) and keep
    # cells with at most 2e5 substeps
    dfloor = tc(1e-30, 1, -3, 0)
    R = tc(R_CGS, 0, 2, -2)
    pfloor = 0.1 * R * dfloor * 0.1
    t = oracle.thermo_params(c, pfloor=pfloor, subcycling=False)
    est = oracle.thermo_batch(t, inp9, 8)
    ratio = inp9[:, 8] / np.abs(0.1 * inp9[:, 4] / est[:, 1])
    keep = np.where(~(ratio > 2e5))[0][:4000]
    inp9 = inp9[keep]
    n = inp9.shape[0]
    # gate cases: ADV below thermo_hii_switch -> exact no-op
    inp9[:100, 3] = 0.005
    # harness input: den, pre, hii, adv, T_min, col_den, photonRate, dt, E.
    # photonRate S such that fluxFUV(0.5 S, r^2) = F_FUV with the harness's
    # r^2 = dx^2 (dx = side_length / 2 in code units).
    side = 1.0
    dx = tc(side, 0, 1, 0) / 2.0
    rsq = ((1.0 * 1.0) * dx) * dx
    S = inp9[:, 7] * (1.2e7 * 4 * 3.14159265359 * rsq) / 0.5
    hin = np.stack([inp9[:, 0], inp9[:, 1], inp9[:, 2], inp9[:, 3], inp9[:, 5], inp9[:, 6], S, inp9[:, 8],
                    inp9[:, 4]], axis=1)
    with tempfile.TemporaryDirectory() as d:
        hin.tofile(f"{d}/in.bin")
        with open(f"{d}/case.txt", "w") as f:
            f.write("density_scale 1e-20\npressure_scale 1e-28\ntime_scale 1e11\n"
                    f"side_length {side}\ngamma 1.67\ndensity_floor 1e-30\ntemperature_floor 0.1\n"
                    "mass_fraction_hydrogen 1.0\nthermo_hii_switch 1e-2\nthermo_subcycling 1\n"
                    f"ncells_in {n}\ninput {d}/in.bin\noutput {d}/out.bin\n")
        subprocess.run([HARNESS, "thermo", f"{d}/case.txt"], check=True)
        out = np.fromfile(f"{d}/out.bin").reshape(n, 10)
    # the harness reports the F_FUV and COL_DEN the reference actually used
    inp9[:, 7] = out[:, 5]
    inp9[:, 6] = out[:, 4]
    np.savez_compressed(f"{HERE}/thermo.npz", inputs=inp9, heat=out[:, 0], rate_pre=out[:, 1], rate=out[:, 2],
                        h_tot=out[:, 3], temperature=out[:, 7], pfloor=pfloor)
    print("thermo", n)


def raytrace():
    """The star-on thermo preTimeStep on small grids: the reference's star
    geometry (Radiation::initField -> Grid::calculateNearestNeighbours,
    cellPathLength, shellVolume), CausalWind cells, the column-density ray
    trace (Thermodynamics::rayTrace) and the pre-step RATE / HEAT.  Random
    clumpy media in code units of the 1e-20/1e-28/1e11 scales; side 1 pc;
    photon rate 1e49 /s."""
    c = oracle.converter(1e-20, 1e-28, 1e11)
    tc = lambda v, m, l, t: oracle.to_code(c, v, m, l, t)  # noqa: E731
    cases = {"a": ((7, 6, 5), (2, 1, 4), 1), "b": ((8, 8, 8), (0, 0, 0), 0), "c": ((9, 10, 11), (7, 3, 2), 2),
             "d": ((6, 6, 6), (5, 5, 5), 0)}
    rng = np.random.default_rng(23)
    out = {}
    for name, (shape, star, wind) in cases.items():
        nz, ny, nx = shape
        rho = 10.0 ** rng.uniform(-24, -20, shape)
        T = 10.0 ** rng.uniform(2, 4, shape)
        hii = rng.uniform(0, 1, shape)
        Q = np.zeros((7,) + shape)
        Q[0] = rho * tc(1.0, 1, -3, 0)
        Q[1] = rho * R_CGS * (hii + 1.0) * T * tc(1.0, 1, -1, -2)
        Q[2] = hii
        Q[3] = np.where(rng.uniform(0, 1, shape) < 0.1, 0.005, 1.0)
        with tempfile.TemporaryDirectory() as d:
            Q.tofile(f"{d}/q.bin")
            kv = {"nx": nx, "ny": ny, "nz": nz, "side_length": 3.0857e18, "density_scale": 1e-20,
                  "pressure_scale": 1e-28, "time_scale": 1e11, "gamma": 1.67, "density_floor": 1e-30,
                  "temperature_floor": 0.1, "mass_fraction_hydrogen": 1.0, "thermo_hii_switch": 1e-2,
                  "star_x": star[0], "star_y": star[1], "star_z": star[2], "wind_radius": wind,
                  "photon_rate": 1e49, "input_q": f"{d}/q.bin", "output": f"{d}/out.bin"}
            with open(f"{d}/case.txt", "w") as f:
                for k, v in kv.items():
                    f.write(f"{k} {v}\n")
            r = subprocess.run([HARNESS, "raytrace", f"{d}/case.txt"], capture_output=True, text=True, check=True)
            o = np.fromfile(f"{d}/out.bin").reshape(nz, ny, nx, 15)
        info = json.loads(r.stdout.strip().splitlines()[-1])
        out[f"{name}_shape"] = np.array(shape)
        out[f"{name}_star"] = np.array(star)
        out[f"{name}_wind_radius"] = np.array(wind)
        out[f"{name}_Q"] = Q
        out[f"{name}_out"] = o
        out[f"{name}_info"] = np.array([info["photon_rate"], info["hydrogen_mass"], info["pi"], info["dx"]])
    np.savez_compressed(f"{HERE}/raytrace.npz", **out)
    print("raytrace", list(cases))


RAD_CASES = {
    # the reference config's radiation settings (config/torch-config.lua):
    # implicit scheme, neq coupling, K1 = 0.2
    "neq": dict(shape=(8, 9, 10), star=(3, 4, 2), wind=1, nsteps=3, side=3.0857e17, T=(1.5, 2.5), hii=0.01,
                kv=dict(K1=0.2, coupling="neq", integration_scheme="implicit", collisions_on=0)),
    # two-temperature isothermal coupling, explicit scheme, collisions, K3/K4
    "tti": dict(shape=(7, 6, 9), star=(0, 0, 0), wind=0, nsteps=2, side=6e17, T=(3.0, 4.3), hii=0.5,
                kv=dict(K3=0.1, K4=0.05, coupling="tti", integration_scheme="explicit", collisions_on=1)),
    # implicit2, no coupling, K1 + K2, heating amplification, wind radius 2
    "off": dict(shape=(9, 7, 12), star=(6, 3, 4), wind=2, nsteps=2, side=1e18, T=(1.0, 3.5), hii=0.2,
                kv=dict(K1=0.5, K2=1.0, coupling="off", integration_scheme="implicit2", collisions_on=1,
                        heating_amplification=2.0)),
}


def radiation():
    """The Radiation integrator (Radiation.cpp) on its own: initField, then
    per step preTimeStepCalculations, calculateTimeStep and two subSteps of
    the radiation component; random clumpy neutral-to-partly-ionised media,
    a star of 1e49 photons/s of 2.976e-11 erg (code units of the
    1e-20/1e-28/1e11 scales).  Stores the inputs, the last step's U, R and
    Q[HII, ADV], and every dt."""
    c = oracle.converter(1e-20, 1e-28, 1e11)
    tc = lambda v, m, l, t: oracle.to_code(c, v, m, l, t)  # noqa: E731
    rng = np.random.default_rng(31)
    out = {}
    for name, cs in RAD_CASES.items():
        shape = cs["shape"]
        nz, ny, nx = shape
        rho = 10.0 ** rng.uniform(-22, -20, shape)
        T = 10.0 ** rng.uniform(cs["T"][0], cs["T"][1], shape)
        hii = rng.uniform(0, cs["hii"], shape)
        Q = np.zeros((7,) + shape)
        Q[0] = rho * tc(1.0, 1, -3, 0)
        Q[1] = rho * R_CGS * (hii + 1.0) * T * tc(1.0, 1, -1, -2)
        Q[2] = hii
        Q[3] = np.where(rng.uniform(0, 1, shape) < 0.5, 0.0, 1.0)
        Q[4:7] = rng.normal(size=(3,) + shape) * 1e5 * tc(1.0, 0, 1, -1)
        with tempfile.TemporaryDirectory() as d:
            Q.tofile(f"{d}/q.bin")
            kv = {"nx": nx, "ny": ny, "nz": nz, "side_length": cs["side"], "density_scale": 1e-20,
                  "pressure_scale": 1e-28, "time_scale": 1e11, "gamma": 1.67, "density_floor": 1e-30,
                  "temperature_floor": 0.1, "mass_fraction_hydrogen": 1.0, "simulation_time": 1e13,
                  "star_x": cs["star"][0], "star_y": cs["star"][1], "star_z": cs["star"][2],
                  "wind_radius": cs["wind"], "photon_rate": 1e49, "photon_energy": 2.976e-11,
                  "photoion_cross_section": 6.3e-18, "case_b_recombination_coeff": 2.59e-13, "tau_0": 0.6,
                  "temperature_hi": 300, "temperature_hii": 10000, "nsteps": cs["nsteps"], "input_q": f"{d}/q.bin",
                  "output_prefix": f"{d}/rad"}
            kv.update(cs["kv"])
            with open(f"{d}/case.txt", "w") as f:
                for k, v in kv.items():
                    f.write(f"{k} {v}\n")
            subprocess.run([HARNESS, "radiation", f"{d}/case.txt"], capture_output=True, text=True, check=True)
            o = np.fromfile(f"{d}/rad_{cs['nsteps']:04d}.bin").reshape((15,) + shape)
            dts = np.fromfile(f"{d}/rad_dt.bin")
        out[f"{name}_Q"] = Q
        out[f"{name}_U"] = o[:7]
        out[f"{name}_R"] = o[7:13]
        out[f"{name}_q"] = o[13:15]
        out[f"{name}_dts"] = dts
    np.savez_compressed(f"{HERE}/radiation.npz", **out)
    print("radiation", list(RAD_CASES))


STAR_FULLSTEP = dict(n=16, star=(8, 7, 9), wind=1, nsteps=4, side=3.0857e18, seed=41,
                     kv={"cooling_on": 1, "radiation_on": 1, "star_on": 1, "photon_rate": 4.9e48,
                         "photon_energy": 2.976e-11, "K1": 0.2, "coupling": "neq", "integration_scheme": "implicit",
                         "photoion_cross_section": 6.3e-18, "case_b_recombination_coeff": 2.59e-13, "tau_0": 0.6,
                         "temperature_hi": 300, "temperature_hii": 10000, "thermo_hii_switch": 1e-2})


def star_fullstep_q():
    """The cgs primitives of the composed star-on case: a clumpy neutral
    medium (1e-22..1e-20 g/cm^3, 100 K, ADV = 1) with small random motions."""
    cs = STAR_FULLSTEP
    n = cs["n"]
    rng = np.random.default_rng(cs["seed"])
    rho = 10.0 ** rng.uniform(-22, -20, (n, n, n))
    Q = np.zeros((7, n, n, n))
    Q[0] = rho
    Q[1] = rho * R_CGS * 100.0
    Q[3] = 1.0
    Q[4:7] = rng.normal(size=(3, n, n, n)) * 1.0e5
    return Q


def star_fullstep():
    """Torch::fullStep with hydrodynamics, thermodynamics and radiation
    around a star (the reference harness's driver loop with all three
    components): 4 steps of a 16^3 clumpy medium; U after every step and
    every dt."""
    cs = STAR_FULLSTEP
    n = cs["n"]
    Q = star_fullstep_q()
    extra = {"density_scale": 1e-20, "pressure_scale": 1e-28, "time_scale": 1e11, "side_length": cs["side"],
             "temperature_floor": 0.1, "simulation_time": 1e13, "ncheckpoints": 10, "star_x": cs["star"][0],
             "star_y": cs["star"][1], "star_z": cs["star"][2], "wind_radius": cs["wind"], **cs["kv"]}
    dumps = [1, 2, cs["nsteps"]]
    out, dts, info = run_hydro_case(Q, cs["nsteps"], dumps, tfloor=0.1, extra=extra)
    np.savez_compressed(f"{HERE}/star_fullstep.npz", Q=Q, dts=dts, time=info["time"],
                        **{f"U_{k}": out[k] for k in dumps})
    print("star_fullstep", info, dts)


MINTEMP = dict(n=12, nsteps=6, side=3.0857e18, seed=43, kv={"cooling_on": 1, "thermo_hii_switch": 1e-2})


def mintemp_q():
    """The cgs primitives of the min_temp_initial_state case: cool (150-250 K),
    nearly neutral gas of ~1e-22 g/cm^3 with ADV = 1, at rest, where the
    soft landing below T_min + 200 K acts when T_min is 100 K."""
    cs = MINTEMP
    n = cs["n"]
    rng = np.random.default_rng(cs["seed"])
    rho = 1.0e-22 * 10.0 ** rng.uniform(-0.5, 0.5, (n, n, n))
    hii = rng.uniform(0.0, 0.05, rho.shape)
    T = rng.uniform(150.0, 250.0, rho.shape)
    Q = np.zeros((7, n, n, n))
    Q[0] = rho
    Q[1] = rho * 8.314462e7 * (1.0 + hii) * T
    Q[2] = hii
    Q[3] = 1.0
    return Q


def mintemp():
    """Thermodynamics with min_temp_initial_state on and off through the
    reference harness's driver loop (hydro + thermo): U after the last step
    and every dt of both runs."""
    cs = MINTEMP
    Q = mintemp_q()
    res = {}
    for on in (1, 0):
        extra = {"density_scale": 1e-20, "pressure_scale": 1e-28, "time_scale": 1e11, "side_length": cs["side"],
                 "temperature_floor": 0.1, "simulation_time": 1e13, "ncheckpoints": 10,
                 "min_temp_initial_state": on, **cs["kv"]}
        out, dts, info = run_hydro_case(Q, cs["nsteps"], [cs["nsteps"]], tfloor=0.1, extra=extra)
        res[on] = (out[cs["nsteps"]], dts)
        print("mintemp", on, info, dts)
    assert not np.array_equal(res[1][0], res[0][0])
    np.savez_compressed(f"{HERE}/mintemp.npz", Q=Q, U_on=res[1][0], dts_on=res[1][1], U_off=res[0][0],
                        dts_off=res[0][1])


DEBUG_STOP = dict(n=16, nsteps=14)


def converging_q(n: int) -> np.ndarray:
    """A converging flow on the unit cube: rho = 1, p = 0.1, v = -2 (x - 1/2):
    the gas heats up in the centre, so the CFL dt rises and then falls."""
    x = (np.arange(n) + 0.5) / n - 0.5
    z, y, xx = np.meshgrid(x, x, x, indexing="ij")
    Q = np.zeros((7, n, n, n))
    Q[0] = 1.0
    Q[1] = 0.1
    Q[4], Q[5], Q[6] = -2.0 * xx, -2.0 * y, -2.0 * z
    return Q


def debug_stop():
    """Torch::calculateTimeStep's debug test (Torch.cpp:321-332: the loop ends
    after the step whose hydro dt has 100 dt / tmax <= 1e-6; Integration.debug
    defaults to true, Parameters.hpp:50) on a 16^3 converging flow:
    - "early": simulation_time 1e6, the stop after step 2 (the first CFL step);
    - "late": simulation_time set between 1e8 times the dt of a later step and
      that of every step before it, so the stop comes after that step;
    - "nodebug": the early case with debug = 0, which runs all its steps;
    - "run": simulation_time 1e3 with debug on, which never stops.
    U after the last step run, every dt, the time and the quitting flag."""
    cs = DEBUG_STOP
    Q = converging_q(cs["n"])
    res = {"Q": Q, "tfloor": 1e-12, "nsteps": cs["nsteps"]}

    def run(tmax, debug, dumps=()):
        return run_hydro_case(Q, cs["nsteps"], list(dumps), tfloor=1e-12,
                              extra={"simulation_time": repr(float(tmax)), "debug": debug})

    # the dts of a run without the test find a step k > 2 whose dt is below
    # every earlier CFL dt
    _, dts0, _ = run(1e6, 0)
    k = next(i for i in range(3, len(dts0)) if dts0[i] < dts0[1:i].min())
    cases = {"early": (1e6, 1), "late": (float(1e8 * 0.5 * (dts0[k] + dts0[1:k].min())), 1), "nodebug": (1e6, 0),
             "run": (1e3, 1)}
    for name, (tmax, debug) in cases.items():
        _, dts, info = run(tmax, debug)
        steps = info["steps"]
        out, dts2, _ = run(tmax, debug, [steps])
        assert np.array_equal(dts, dts2)
        res[f"{name}_tmax"] = tmax
        res[f"{name}_debug"] = debug
        res[f"{name}_dts"] = dts
        res[f"{name}_time"] = info["time"]
        res[f"{name}_steps"] = steps
        res[f"{name}_quitting"] = info["quitting"]
        res[f"{name}_U"] = out[steps]
        print("debug_stop", name, tmax, info)
    assert res["early_steps"] == 2 and res["early_quitting"] == 1
    assert res["late_steps"] == k + 1 and res["late_quitting"] == 1
    assert res["nodebug_steps"] == cs["nsteps"] and res["nodebug_quitting"] == 0
    assert res["run_steps"] == cs["nsteps"] and res["run_quitting"] == 0
    np.savez_compressed(f"{HERE}/debug_stop.npz", **res)


CLOUD_SPLIT = dict(n=16, nsteps=6, seed=None)


def cloud_split():
    """Torch::fullStep with hydrodynamics and thermodynamics, star off, on a
    16^3 version of the operator-split cloud (torch_amd.cloud.setup in cgs:
    side 0.5 pc, n(R) = n0 / (1 + R^2 / rc^2) about a centre 0.35 pc from
    the grid centre, T = 300 K, HII fraction 1 within 0.1 pc of the centre,
    ADV = 1), scales 1e-20 / 1e-28 / 1e11, temperature_floor 0.1, the star
    cell (the CausalWind cell the reference keeps with the star off) at the
    grid centre.  With the star off the reference's cooling sees no FUV flux
    and zero column density.  U after every step and every dt."""
    from torch_amd import cloud

    cs = CLOUD_SPLIT
    n = cs["n"]
    Q, _, _, _, side = cloud.setup(n, lambda v, m, l, t: v)
    c = n // 2
    extra = {"density_scale": 1e-20, "pressure_scale": 1e-28, "time_scale": 1e11, "side_length": repr(side),
             "temperature_floor": 0.1, "simulation_time": 1e13, "ncheckpoints": 10, "cooling_on": 1,
             "thermo_hii_switch": 1e-2, "star_x": c, "star_y": c, "star_z": c}
    dumps = list(range(1, cs["nsteps"] + 1))
    out, dts, info = run_hydro_case(Q, cs["nsteps"], dumps, tfloor=0.1, extra=extra)
    np.savez_compressed(f"{HERE}/cloud_split.npz", Q=Q, side=side, dts=dts, time=info["time"],
                        **{f"U_{k}": out[k] for k in dumps})
    print("cloud_split", info, dts)


def heating():
    """Thermodynamics::fillHeatingArrays as Torch::run calls it at a
    checkpoint, through the reference harness's loop: GridCell::H and Fluid::Q
    after the last step, then H after each of two fills (TOT accumulates
    RHII + EUVH until an integrate resets it).  Cases: the composed star run
    (hydro, cooling, radiation; 4 steps), the star-off cloud (hydro, cooling;
    6 steps) and the star run without cooling (hydro, radiation; TOT is never
    reset)."""
    cases = {}
    cs = STAR_FULLSTEP
    base_star = {"density_scale": 1e-20, "pressure_scale": 1e-28, "time_scale": 1e11, "side_length": cs["side"],
                 "temperature_floor": 0.1, "simulation_time": 1e13, "ncheckpoints": 10, "star_x": cs["star"][0],
                 "star_y": cs["star"][1], "star_z": cs["star"][2], "wind_radius": cs["wind"], **cs["kv"]}
    cases["star"] = (star_fullstep_q(), cs["nsteps"], base_star)
    cases["radonly"] = (star_fullstep_q(), 3, dict(base_star, cooling_on=0))
    from torch_amd import cloud

    cl = CLOUD_SPLIT
    Qc, _, _, _, side = cloud.setup(cl["n"], lambda v, m, l, t: v)
    c = cl["n"] // 2
    cases["cloud"] = (Qc, cl["nsteps"], {"density_scale": 1e-20, "pressure_scale": 1e-28, "time_scale": 1e11,
                                         "side_length": repr(side), "temperature_floor": 0.1,
                                         "simulation_time": 1e13, "ncheckpoints": 10, "cooling_on": 1,
                                         "thermo_hii_switch": 1e-2, "star_x": c, "star_y": c, "star_z": c})
    res = {}
    for name, (Q, nsteps, extra) in cases.items():
        nz, ny, nx = Q.shape[1:]
        with tempfile.TemporaryDirectory() as d:
            Q.astype(np.float64).tofile(f"{d}/q.bin")
            kv = {"nx": nx, "ny": ny, "nz": nz, "gamma": 1.67, "density_floor": 1e-30, "nsteps": nsteps,
                  "input_q": f"{d}/q.bin", "dump_steps": "", "output_prefix": f"{d}/ref", "heating_fills": 2}
            kv.update(extra)
            with open(f"{d}/case.txt", "w") as f:
                for k, v in kv.items():
                    f.write(f"{k} {v}\n")
            r = subprocess.run([HARNESS, "hydro", f"{d}/case.txt"], capture_output=True, text=True, check=True)
            info = json.loads(r.stdout.strip().splitlines()[-1])
            res[f"{name}_Q0"] = Q
            res[f"{name}_nsteps"] = nsteps
            res[f"{name}_dts"] = np.fromfile(f"{d}/ref_dt.bin")
            res[f"{name}_time"] = info["time"]
            res[f"{name}_Qprint"] = np.fromfile(f"{d}/ref_Q.bin").reshape(7, nz, ny, nx)
            for k in range(3):
                res[f"{name}_H{k}"] = np.fromfile(f"{d}/ref_H_{k}.bin").reshape(11, nz, ny, nx)
        print("heating", name, info)
    # the accumulation is visible: without cooling, TOT grows by RHII + EUVH
    tot1, tot2 = res["radonly_H1"][10], res["radonly_H2"][10]
    assert not np.array_equal(tot1, tot2)
    np.savez_compressed(f"{HERE}/heating.npz", **res)


if __name__ == "__main__":
    if not os.path.exists(HARNESS):
        sys.exit("build the reference harness first: make -C oracle ref")
    which = sys.argv[1:] or ["consts", "riemann", "thermo", "variants", "turbulent", "sod"]
    for w in which:
        globals()[w]()
