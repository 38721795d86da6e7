"""Timing of the star-on operators on one GPU: the thermo column-density ray
trace (Thermodynamics::rayTrace inside the star-on thermo preTimeStep) and
one radiation sub-step (Radiation::preTimeStepCalculations + integrate +
updateSourceTerms, implicit scheme, neq coupling), on an n^3 clumpy neutral
medium (n_H ~ 1e2..1e4 cm^-3, 100 K) with the star at the centre, wind
radius 2, side 1 pc, 4.9e48 photons/s.

    python -m torch_amd.star_bench [--n 256] [--reps 5] [--fullstep N]

Prints one JSON line: GPU seconds per call and Mcell/s of each operator
(HIP events around `reps` calls after one warm-up call; inputs resident in
HBM).  bench.py adds the CPU baseline of the same operators.
"""
from __future__ import annotations

import argparse
import json

import numpy as np

SCALES = (1e-20, 1e-28, 1e11)


def medium(n: int, units):
    """Primitives [7][n][n][n] in code units and the cgs side length."""
    from torch_amd import setups

    Qt = setups.turbulent(n, 1.0, 0.5)
    rho = 1.674e-24 * 1.0e3 * Qt[0]
    Q = np.zeros((7, n, n, n))
    Q[0] = rho * units(1.0, 1, -3, 0)
    Q[1] = rho * 8.314462e7 * 100.0 * units(1.0, 1, -1, -2)
    Q[3] = 1.0
    return Q, 3.0857e18


def gpu_rates(n: int = 256, reps: int = 5) -> dict:
    """GPU seconds per call and Mcell/s of the ray trace and of a radiation
    sub-step on the n^3 star workload."""
    import torch

    import torch_amd as ta

    units = lambda v, m, l, t: ta.units(SCALES, v, m, l, t)  # noqa: E731
    Q, side = medium(n, units)
    R = units(8.314462e7, 0, 2, -2)
    dfloor = units(1e-30, 1, -3, 0)
    p = ta.make_params((n, n, n), side_length=units(side, 0, 1, 0), gamma=1.67, dfloor=dfloor,
                       pfloor=0.1 * R * dfloor * 0.1, tfloor=0.1, spec_gas_const=R, tmax=units(1e13, 0, 0, 1))
    g = ta.Grid(p)
    g.set_primitive(Q)
    star = ta.make_star((n // 2, n // 2, n // 2), 2, units(4.9e48, 0, 0, -1))
    cool = ta.Cooling(ta.thermo_params(scales=SCALES))
    cool.set_star(g, star)
    rad = ta.Radiation(ta.rad_params(SCALES, K=(0.2, 0, 0, 0), coupling="neq"), g, star)
    rad.init_field(Q)
    stream = torch.cuda.current_stream()

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / 1e3 / reps

    # the thermo timestep with and without the star: the difference is the
    # ray trace (both form their inputs from the grid in the rate kernel)
    t_with = timed(lambda: cool.grid_timestep(g, p.dt_max))
    cool.set_star(g, None)
    t_without = timed(lambda: cool.grid_timestep(g, p.dt_max))
    cool.set_star(g, star)
    rad.pre_step()
    dt = rad.timestep(p.dt_max)
    t_rad = timed(lambda: (rad.pre_step(), rad.substep(dt, True)))
    g.synchronize()
    return {"n": n, "cells": n ** 3, "reps": reps, "shells": n // 2 + 1,
            "thermo_timestep_star_s": t_with, "thermo_timestep_nostar_s": t_without,
            "ray_trace_s": t_with - t_without, "ray_trace_mcells_per_s": n ** 3 / (t_with - t_without) / 1e6,
            "radiation_dt": dt, "radiation_substep_s": t_rad, "radiation_mcells_per_s": n ** 3 / t_rad / 1e6}


def fullstep_rates(n: int = 128, steps: int = 5, warm: int = 2) -> dict:
    """ms per Torch::fullStep (hydro FP64, star-on cooling, radiation) on the
    n^3 star workload: the host-driven FullStep (one host round trip per
    time-step reduction) and the device loop (torchgpu_run_full), each after
    `warm` steps, on two grids that follow the same trajectory."""
    import time

    import torch

    import torch_amd as ta
    from torch_amd import fullstep

    units = lambda v, m, l, t: ta.units(SCALES, v, m, l, t)  # noqa: E731
    Q, side = medium(n, units)
    R = units(8.314462e7, 0, 2, -2)
    dfloor = units(1e-30, 1, -3, 0)
    out = {"n": n, "steps": steps, "warm": warm}
    for mode in ("host", "device"):
        p = ta.make_params((n, n, n), side_length=units(side, 0, 1, 0), gamma=1.67, dfloor=dfloor,
                           pfloor=0.1 * R * dfloor * 0.1, tfloor=0.1, spec_gas_const=R, tmax=units(1e13, 0, 0, 1))
        g = ta.Grid(p)
        g.set_primitive(Q)
        star = ta.make_star((n // 2, n // 2, n // 2), 2, units(4.9e48, 0, 0, -1))
        tp = ta.thermo_params(scales=SCALES)
        tp.pfloor = p.pfloor
        cool = ta.Cooling(tp)
        cool.set_star(g, star)
        rad = ta.Radiation(ta.rad_params(SCALES, K=(0.2, 0, 0, 0), coupling="neq"), g, star)
        rad.init_field(Q)
        loop = fullstep.FullStep(g, cooling=cool, radiation=rad)
        run = (lambda k: [loop.step() for _ in range(k)]) if mode == "host" else \
            (lambda k: rad.run_full(k, cooling=cool))
        run(warm)
        g.synchronize()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        run(steps)
        g.synchronize()
        out[f"{mode}_ms_per_step"] = (time.perf_counter() - t0) * 1e3 / steps
        out[f"{mode}_time"] = g.clock()[0] if mode == "device" else loop.time
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--fullstep", type=int, default=0, metavar="STEPS",
                    help="time STEPS full star steps, host-driven and device loop, instead")
    a = ap.parse_args()
    if a.fullstep:
        print(json.dumps(fullstep_rates(a.n, a.fullstep)), flush=True)
        return
    print(json.dumps(gpu_rates(a.n, a.reps)), flush=True)


if __name__ == "__main__":
    main()
