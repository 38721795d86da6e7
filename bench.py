#!/usr/bin/env python3
"""bench.py -- hydro step throughput on one MI355X.

Headline workload (the 512^3 mixed-precision case): 512^3 cells on the unit
cube, gamma = 1.67, outflow boundaries, density log-normal (sigma_ln rho =
1.5, mean 1) from a seeded Gaussian random field with P(k) ~ k^-11/3 on
2 <= |k| <= 32, uniform pressure 1, velocities from three more such fields
scaled to rms Mach 4; 2nd-order (linear reconstruction, van Albada) and the
rotated HLLC-HLL solver; Torch::run step sequence (first step dt_max*1e-20,
then CFL 0.5); Integration.debug off, as the reference's shipped
configuration has it.

One "step" is one Torch::hydroStep (predictor + corrector) of the whole grid
with the CFL reduction for the next step; a cell-update is one cell advanced
by one step.  `value` is the mixed-precision path (FP32 reconstruction and
Riemann fluxes, FP64 accumulation, pressure recovery, floors and CFL); the
FP64 path's rate, the speed-up and the mixed path's density, pressure and
velocity error against FP64 after 200 steps are reported alongside (both
runs take the smaller of their two CFL limits each step).  Inputs are
resident in HBM before the timed region.  Extra keys: the 256^3 box
(sigma 1, Mach 2), the cooling step on 2^24 independent cells (FP64
subcycled integrator vs the surrogate step), the operator-split cloud
(384^3, 100 steps, surrogate cooling with the network trained on the
cloud's conditions) and the star operators.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 512]
"""
from __future__ import annotations

import argparse
import json
import os
import shutil
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
SIGMA, MACH, GAMMA = 1.5, 4.0, 1.67     # the 512^3 mixed-precision workload
SIGMA2, MACH2 = 1.0, 2.0                # the 256^3 box
TFLOOR = 1e-12  # with the C++ default 0 the reference's step aborts on a zero pressure
R_CGS = 8.314462e7
PFLOOR = 0.1 * R_CGS * 1e-30 * TFLOOR
# algorithmic HBM bytes per cell of a mixed-precision stage: 7 FP64
# conserved variables (56 B) and their 7 float primitives (28 B); the
# corrector reads Q(U*) and U^n and writes U^n+1 and Q(U^n+1)
BYTES_STAGE2 = 28 + 56 + 56 + 28


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(n_sample: int, steps: int, threads: int) -> dict:
    """The reference's FP64 CPU step on the host cores (before the GPU is
    touched): the reference harness under mpirun when it is present, else
    the oracle restatement with OpenMP threads.  Bounded sample: a
    n_sample^3 box of the same turbulent set-up for `steps` steps."""
    import numpy as np

    from torch_amd import setups

    Q = setups.turbulent(n_sample, SIGMA, MACH, gamma=GAMMA)
    harness = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
    mpirun = shutil.which("mpirun") or ("/opt/conda/bin/mpirun" if os.path.exists("/opt/conda/bin/mpirun") else None)
    if os.path.exists(harness) and mpirun:
        try:
            with tempfile.TemporaryDirectory() as d:
                Q.tofile(f"{d}/q.bin")
                with open(f"{d}/case.txt", "w") as f:
                    f.write(f"nx {n_sample}\nny {n_sample}\nnz {n_sample}\nside_length 1.0\ngamma {GAMMA}\n"
                            f"density_floor 1e-30\ntemperature_floor {TFLOOR}\nsimulation_time 1000\n"
                            f"ncheckpoints 1\ndebug 0\nnsteps {steps}\ninput_q {d}/q.bin\n")
                nproc = min(threads, n_sample)
                r = subprocess.run([mpirun, "-np", str(nproc), harness, "hydro", f"{d}/case.txt"],
                                   capture_output=True, text=True, timeout=300)
                if r.returncode == 0:
                    info = json.loads(r.stdout.strip().splitlines()[-1])
                    return {"value": round(info["mcell_updates_per_s"], 4), "unit": "Mcell-updates/s",
                            "cores": nproc, "kind": "reference",
                            "sample": f"{n_sample}^3 turbulent box (sigma_ln_rho {SIGMA}, Mach {MACH}), {info['steps']} steps, FP64, "
                                      f"mpirun -np {nproc} (x-slab decomposition)"}
                log("reference harness failed:", r.stderr[-500:])
        except Exception as e:  # noqa: BLE001
            log("reference harness unavailable:", e)
    import oracle

    p = oracle.make_params(Q.shape[1:], 1.0 / n_sample, gamma=GAMMA, tfloor=TFLOOR, pfloor=PFLOOR,
                           nthreads=threads)
    U0 = oracle.UfromQ_grid(oracle.fix_primitives(p, Q, np.zeros_like(Q)), GAMMA)
    t0 = time.perf_counter()
    oracle.run_hydro(p, U0, steps, debug=False)
    dt = time.perf_counter() - t0
    return {"value": round(n_sample ** 3 * steps / dt / 1e6, 4), "unit": "Mcell-updates/s", "cores": threads,
            "kind": "port", "sample": f"{n_sample}^3 turbulent box (sigma_ln_rho {SIGMA}, Mach {MACH}), {steps} steps, FP64 oracle, {threads} OpenMP threads"}


def measured_traffic(n: int, kernel: str):
    """HBM bytes per launch of `kernel` (FETCH_SIZE + WRITE_SIZE, scaled by a
    device copy of known size) from the committed rocprofv3 summary of the
    same workload; None when no summary covers this grid size."""
    path = os.path.join(ROOT, "profiles", f"hydro_stage_turb{n}_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        d = json.load(fh)
    k = d.get("kernels", {}).get(kernel)
    return None if k is None else round(k["read_bytes"] + k["write_bytes"])


def _event_pair():
    import torch

    return torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)


def hydro_rates(Q, n: int, steps: int, warmup: int):
    """FP64 and mixed-precision Torch::run steps on the field Q (device
    tensor): Mcell-updates/s over `steps` device-loop steps after `warmup`
    untimed ones, and the two stage kernels' launch times (HIP events on the
    stream the library enqueues on)."""
    import torch

    import torch_amd as ta

    stream = torch.cuda.current_stream()
    res = {}
    for name, prec in (("fp64", ta.FP64), ("mixed", ta.MIXED)):
        g = ta.Grid(ta.make_params((n, n, n), gamma=GAMMA, tfloor=TFLOOR, pfloor=PFLOOR, precision=prec, debug=False),
                    stream=stream.cuda_stream)
        g.set_primitive(Q)
        g.run(warmup)
        g.synchronize()
        torch.cuda.synchronize()
        e0, e1 = _event_pair()
        e0.record(stream)
        g.run(steps)
        e1.record(stream)
        torch.cuda.synchronize()
        g.synchronize()
        ms = e0.elapsed_time(e1) / steps
        # per-kernel timing of the two stage kernels (same dt as the last step)
        _, last_dt, _ = g.clock()
        g.set_dt(last_dt)
        reps = max(3, min(10, steps))
        ev = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(reps)]
        for r in range(reps):
            ev[r][0].record(stream)
            g.stage(1)
            ev[r][1].record(stream)
            g.stage(2)
            ev[r][2].record(stream)
        torch.cuda.synchronize()
        g.synchronize()
        s1 = sum(e[0].elapsed_time(e[1]) for e in ev) / reps
        s2 = sum(e[1].elapsed_time(e[2]) for e in ev) / reps
        res[name] = {"ms_per_step": ms, "mcells": n ** 3 / (ms * 1e-3) / 1e6, "stage1_ms": s1, "stage2_ms": s2}
        log(f"{n}^3 {name}: {ms:.3f} ms/step  {res[name]['mcells']:.1f} Mcell-updates/s  stage1 {s1:.3f} ms  "
            f"stage2 {s2:.3f} ms")
        g.close()
    torch.cuda.empty_cache()
    return res


def mixed_accuracy(Q, n: int, steps: int, every: int = 50):
    """Density, pressure and velocity L1 / L-inf error of the mixed path
    against the FP64 path after `steps` steps from the same field.  Both runs
    take dt = min(their two CFL limits) each step (the first step
    dt_max*1e-20): with the FP64 run's dt alone the mixed run can exceed its
    own CFL limit in the near-vacuum rarefactions this field develops."""
    import torch

    import torch_amd as ta

    stream = torch.cuda.current_stream()
    kw = dict(gamma=GAMMA, tfloor=TFLOOR, pfloor=PFLOOR, debug=False)
    g64 = ta.Grid(ta.make_params((n, n, n), precision=ta.FP64, **kw), stream=stream.cuda_stream)
    gmx = ta.Grid(ta.make_params((n, n, n), precision=ta.MIXED, **kw), stream=stream.cuda_stream)
    g64.set_primitive(Q)
    gmx.set_primitive(Q)
    dt_max = g64.params.dt_max
    t0 = time.perf_counter()
    sim_t = 0.0
    for s in range(1, steps + 1):
        dt = dt_max * 1.0e-20 if s == 1 else min(g64.hydro_timestep(dt_max), gmx.hydro_timestep(dt_max))
        g64.hydro_step(dt)
        gmx.hydro_step(dt)
        sim_t += dt
        if s % every == 0:
            g64.synchronize()
            gmx.synchronize()
            log(f"accuracy run: step {s}, t {sim_t:.4g}, dt {dt:.4g} ({time.perf_counter() - t0:.1f}s)")
    g64.synchronize()
    gmx.synchronize()
    q64 = torch.empty((7, n, n, n), dtype=torch.float64, device="cuda")
    qmx = torch.empty_like(q64)
    g64.get_primitive(fix=False, out=q64)
    gmx.get_primitive(fix=False, out=qmx)
    g64.close()
    gmx.close()
    errs = {}
    for key, idx in (("density", [0]), ("pressure", [1]), ("velocity", [4, 5, 6])):
        a, b = q64[idx], qmx[idx]
        d = (b - a).abs()
        errs[key] = {"rel_l1": float(d.sum() / a.abs().sum()), "rel_linf": float(d.max() / a.abs().max())}
    # the velocity of near-vacuum cells (the rarefactions reach densities of
    # 1e-14 of the mean) is momentum over a vanishing density, so its L-inf
    # sits in a handful of such cells; stated also for the cells at >= 1e-6 of
    # the mean density, and as the momentum rho v
    dense = (q64[0] >= 1.0e-6 * q64[0].mean()).unsqueeze(0)
    a, b = q64[4:7], qmx[4:7]
    d = (b - a).abs() * dense
    errs["velocity_dense_cells"] = {"rel_l1": float(d.sum() / (a.abs() * dense).sum()),
                                    "rel_linf": float(d.max() / (a.abs() * dense).max()),
                                    "cells_excluded": int((~dense).sum())}
    ma, mb = q64[0:1] * a, qmx[0:1] * b
    d = (mb - ma).abs()
    errs["momentum"] = {"rel_l1": float(d.sum() / ma.abs().sum()), "rel_linf": float(d.max() / ma.abs().max())}
    del q64, qmx, a, b, d, ma, mb, dense
    torch.cuda.empty_cache()
    return {"steps": steps, "time": sim_t, "dt_rule": "min of both CFL limits", **errs}


def turbulent_field(n: int, sigma: float, mach: float):
    import torch

    from torch_amd import setups

    t0 = time.perf_counter()
    Q = torch.from_numpy(setups.turbulent(n, sigma, mach, gamma=GAMMA, device="cuda")).cuda().contiguous()
    log(f"IC {n}^3 (sigma {sigma}, Mach {mach}) in {time.perf_counter() - t0:.1f}s")
    return Q


def cooling_rates(log2_cells: int, weights: str, gate: float, tail_frac: float = 1e-3) -> dict:
    """The cooling source step on 2^log2_cells independent cells of the
    cooling workload (evaluation seed 11, disjoint from the training seeds):
    the FP64 subcycled integrator and the surrogate step (FP64 prepass, BF16
    MLP, FP64 fallback) on the same inputs, each event-timed on the stream
    it runs on, with the surrogate's relative temperature error against the
    FP64 result.  The FP64 time is set by the longest cells (one cell needs
    ~1e8 serial subcycles), so both steps are also timed on the same cells
    without the `tail_frac` longest: a throughput-bound comparison."""
    import torch

    import torch_amd as ta
    from torch_amd import surrogate as sg
    from torch_amd.train_surrogate import workload_inputs

    cool = ta.Cooling(ta.thermo_params())
    have = os.path.exists(weights)
    mlp = sg.attach(cool, weights, gate=gate) if have else None
    warm = workload_inputs(1 << 12, seed=12)
    cool.thermo_cells(warm)
    if have:
        sg.thermo_cells_surrogate(cool, warm)
    torch.cuda.synchronize()
    n = 1 << log2_cells
    inp = workload_inputs(n, seed=11)
    if not have:
        s = torch.cuda.current_stream()
        e0, e1 = _event_pair()
        e0.record(s)
        _, nsub = cool.thermo_cells(inp)
        e1.record(s)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1)
        return {"cells": n, "fp64_subcycled": {"ms": round(ms, 2), "Mcells_per_s": round(n / ms / 1e3, 4),
                                               "subcycles": int(nsub)}, "surrogate": None}
    r, ref = sg.compare(cool, inp, return_fp64=True)
    # the same cells without the longest tail_frac of them
    nsub = ref[3]
    k = int(n * (1.0 - tail_frac))
    keep = torch.argsort(nsub)[:k]
    cut = float(nsub[keep[-1]])
    sub = inp[:, keep].contiguous()
    del ref, nsub, keep
    rt = sg.compare(cool, sub)
    del sub
    torch.cuda.empty_cache()
    del mlp
    return {
        "cells": n,
        "fp64_subcycled": {"ms": round(r["fp64_ms"], 2), "Mcells_per_s": round(n / r["fp64_ms"] / 1e3, 4),
                           "subcycles": r["fp64_subcycles"],
                           "subcycles_per_s": round(r["fp64_subcycles"] / (r["fp64_ms"] * 1e-3), 1)},
        "surrogate": {"ms": round(r["surrogate_ms"], 2), "Mcells_per_s": round(n / r["surrogate_ms"] / 1e3, 4),
                      "gate": gate, "cells_exact": r["exact"], "cells_mlp": r["mlp_accepted"],
                      "cells_fallback": r["fallback"]},
        "speedup": round(r["fp64_ms"] / r["surrogate_ms"], 3),
        "rel_T_err_all_cells": r["rel_T_err_all"],
        "rel_T_err_mlp_cells": r["rel_T_err_mlp_cells"],
        "without_longest": {
            "cells": k, "excluded_fraction": tail_frac, "max_subcycles_kept": cut,
            "fp64_ms": round(rt["fp64_ms"], 2), "fp64_subcycles": rt["fp64_subcycles"],
            "fp64_subcycles_per_s": round(rt["fp64_subcycles"] / (rt["fp64_ms"] * 1e-3), 1),
            "surrogate_ms": round(rt["surrogate_ms"], 2), "speedup": round(rt["fp64_ms"] / rt["surrogate_ms"], 3),
            "cells_exact": rt["exact"], "cells_mlp": rt["mlp_accepted"], "cells_fallback": rt["fallback"],
            "rel_T_err_mlp_cells": rt["rel_T_err_mlp_cells"]},
    }


def cloud_rates(n: int, steps: int, weights: str, gate: float) -> dict:
    """The operator-split cloud (torch_amd.cloud: 0.5 pc box, the softened
    cloud profile, reference unit scales): `steps` Torch::fullStep
    iterations of mixed-precision hydro sub-steps and the surrogate cooling
    step in the device loop (torchgpu_run_split), event-timed, then the
    cooling step on the grid's final cells compared with the FP64
    integrator at the loop's dt and at 10x, 100x and 1000x it, with the
    cells on each path.  `weights` is the network trained on the cloud's
    conditions (train_surrogate --workload cloud)."""
    import torch

    from torch_amd import cloud
    from torch_amd import surrogate as sg

    t0 = time.perf_counter()
    p, g, cool, thermo = cloud.build(n, "mixed")
    mlp = sg.attach(cool, weights, gate=gate)
    cool.run_split(g, 2, surrogate=True, **thermo)
    g.synchronize()
    log(f"cloud {n}^3 set up in {time.perf_counter() - t0:.1f}s")
    s = torch.cuda.current_stream()
    e0, e1 = _event_pair()
    e0.record(s)
    cool.run_split(g, steps, surrogate=True, **thermo)
    e1.record(s)
    torch.cuda.synchronize()
    g.synchronize()
    ms = e0.elapsed_time(e1) / steps
    t, dt, nst = g.clock()
    out = {"n": n, "steps": steps, "ms_per_step": round(ms, 3),
           "Mcell_updates_per_s": round(n ** 3 / (ms * 1e-3) / 1e6, 2), "time": t, "dt": dt,
           "network": os.path.basename(weights)}
    # the cooling step on the final cells at the loop's dt, and at 10x, 100x
    # and 1000x it (beyond what Thermodynamics::calculateTimeStep lets the
    # loop take), with the cells on each path
    for mult in (1, 10, 100, 1000):
        inp = cloud.cooling_inputs(g, p, thermo["col_den"], thermo["F_FUV"], thermo["excluded"], dt * mult)
        r = sg.compare(cool, inp)
        del inp
        torch.cuda.empty_cache()
        out["cooling_check" if mult == 1 else f"cooling_check_dt_x{mult}"] = {
            "fp64_ms": round(r["fp64_ms"], 2), "surrogate_ms": round(r["surrogate_ms"], 2),
            "fp64_subcycles": r["fp64_subcycles"],
            "cells_exact": r["exact"], "cells_mlp": r["mlp_accepted"], "cells_fallback": r["fallback"],
            "cells_out_of_envelope": r["cells"] - r["exact"] - r["mlp_queued"],
            "rel_T_err_all": r["rel_T_err_all"], "rel_T_err_mlp_cells": r["rel_T_err_mlp_cells"]}
    del mlp
    g.close()
    torch.cuda.empty_cache()
    return out


def star_cpu_baseline(m: int) -> dict:
    """The star operators on the host on an m^3 grid of the same medium: the
    reference's own Thermodynamics::rayTrace and one Radiation sub-step
    (implicit, neq; half of a two-sub-step step) through the reference
    harness when it is present, else the oracle's single-threaded C
    restatement of the same."""
    import numpy as np

    from torch_amd import star_bench

    harness = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
    if os.path.exists(harness):
        try:
            import torch_amd as ta

            units = lambda v, mi, li, ti: ta.units(star_bench.SCALES, v, mi, li, ti)  # noqa: E731
            Q, side = star_bench.medium(m, units)
            common = {"nx": m, "ny": m, "nz": m, "side_length": side, "density_scale": 1e-20, "pressure_scale": 1e-28,
                      "time_scale": 1e11, "gamma": 1.67, "density_floor": 1e-30, "temperature_floor": 0.1,
                      "mass_fraction_hydrogen": 1.0, "simulation_time": 1e13, "star_x": m // 2, "star_y": m // 2,
                      "star_z": m // 2, "wind_radius": 2, "photon_rate": 4.9e48}
            rad = {"photon_energy": 2.976e-11, "K1": 0.2, "coupling": "neq", "integration_scheme": "implicit",
                   "photoion_cross_section": 6.3e-18, "case_b_recombination_coeff": 2.59e-13, "tau_0": 0.6,
                   "temperature_hi": 300, "temperature_hii": 10000, "nsteps": 1}
            with tempfile.TemporaryDirectory() as d:
                Q.tofile(f"{d}/q.bin")
                res = {}
                for mode, extra in (("raytrace", {"output": f"{d}/rt.bin"}), ("radiation", dict(rad, output_prefix=f"{d}/rad"))):
                    kv = dict(common, input_q=f"{d}/q.bin", **extra)
                    with open(f"{d}/{mode}.txt", "w") as f:
                        for k, v in kv.items():
                            f.write(f"{k} {v}\n")
                    r = subprocess.run([harness, mode, f"{d}/{mode}.txt"], capture_output=True, text=True, timeout=300)
                    if r.returncode != 0:
                        raise RuntimeError(r.stderr[-500:])
                    res[mode] = json.loads(r.stdout.strip().splitlines()[-1])
            return {"n": m, "cores": 1, "kind": "reference",
                    "ray_trace_mcells_per_s": round(m ** 3 / res["raytrace"]["ray_trace_s"] / 1e6, 3),
                    "radiation_mcells_per_s": round(m ** 3 / res["radiation"]["substep_s"] / 1e6, 3),
                    "sample": f"{m}^3 star workload through the reference's own units (one process): "
                              "Thermodynamics::rayTrace, one Radiation sub-step"}
        except Exception as e:  # noqa: BLE001
            log("reference harness unavailable for the star baseline:", e)
    import oracle

    c = oracle.converter(*star_bench.SCALES)
    units = lambda v, mi, li, ti: oracle.to_code(c, v, mi, li, ti)  # noqa: E731
    Q, side = star_bench.medium(m, units)
    dx = units(side, 0, 1, 0) / m
    st = (m // 2, m // 2, m // 2)
    geom = oracle.star_geometry((m, m, m), st, dx, 3.14159265359)
    t0 = time.perf_counter()
    oracle.thermo_ray_trace((m, m, m), st, geom, units(1.674e-24, 1, 0, 0), Q[0])
    t_rt = time.perf_counter() - t0
    R = units(8.314462e7, 0, 2, -2)
    dfloor = units(1e-30, 1, -3, 0)
    p = oracle.make_params((m, m, m), dx, gamma=1.67, dfloor=dfloor, pfloor=0.1 * R * dfloor * 0.1, tfloor=0.1,
                           spec_gas_const=R, dt_max=units(1e13, 0, 0, 1) / 10.0, nthreads=1)
    r = oracle.rad_params(c, K=(0.2, 0, 0, 0), coupling="neq", photon_energy=2.976e-11, photon_rate=4.9e48)
    Qf = oracle.fix_primitives(p, Q, np.zeros_like(Q))
    U0 = oracle.UfromQ_grid(Qf, p.gamma)
    R0 = oracle.rad_init_field(r, (m, m, m), st, dx, Qf)
    t0 = time.perf_counter()
    oracle.rad_run(r, p, st, 2, U0, R0, 1)
    t_rad = (time.perf_counter() - t0) / 2.0
    return {"n": m, "cores": 1, "kind": "port", "ray_trace_mcells_per_s": round(m ** 3 / t_rt / 1e6, 3),
            "radiation_mcells_per_s": round(m ** 3 / t_rad / 1e6, 3),
            "sample": f"{m}^3 star workload, oracle restatement: one ray trace, one radiation step of two sub-steps"}


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--accuracy-steps", type=int, default=200, help="steps of the mixed-vs-FP64 error run (0: none)")
    ap.add_argument("--n2", type=int, default=256, help="the sigma 1 / Mach 2 box (0: none)")
    ap.add_argument("--cpu-threads", type=int, default=min(16, os.cpu_count() or 1))
    ap.add_argument("--cpu-sample", type=int, default=96)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cooling-log2", type=int, default=24, help="cooling workload size, 2^k cells")
    ap.add_argument("--no-cooling", action="store_true")
    ap.add_argument("--surrogate", default=os.path.join(ROOT, "torch_amd", "data", "surrogate_cooling.npz"))
    ap.add_argument("--cloud-surrogate", default=os.path.join(ROOT, "torch_amd", "data", "surrogate_cloud.npz"),
                    help="network trained on the cloud's conditions (train_surrogate --workload cloud)")
    ap.add_argument("--gate", type=float, default=0.01)
    ap.add_argument("--cloud-n", type=int, default=384, help="operator-split cloud grid (0: none)")
    ap.add_argument("--cloud-steps", type=int, default=100)
    ap.add_argument("--star-n", type=int, default=256, help="star workload (ray trace, radiation) grid")
    ap.add_argument("--no-star", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for {args.gpus}); "
              "the grid would shard as 3-D blocks with 2-deep halo exchanges per stage.", file=sys.stderr)
        sys.exit(2)

    cpu = None
    if not args.no_cpu:
        t0 = time.perf_counter()
        cpu = cpu_baseline(args.cpu_sample, 3, args.cpu_threads)
        log(f"cpu baseline {cpu} ({time.perf_counter() - t0:.1f}s)")

    import torch

    n = args.n
    Q = turbulent_field(n, SIGMA, MACH)
    res = hydro_rates(Q, n, args.steps, args.warmup)
    errs = None
    if args.accuracy_steps > 0:
        t0 = time.perf_counter()
        errs = mixed_accuracy(Q, n, args.accuracy_steps)
        log(f"mixed vs FP64 after {args.accuracy_steps} steps: {errs} ({time.perf_counter() - t0:.1f}s)")
    del Q
    torch.cuda.empty_cache()
    mx, f64 = res["mixed"], res["fp64"]
    box2 = None
    if args.n2 > 0:
        Q2 = turbulent_field(args.n2, SIGMA2, MACH2)
        r2 = hydro_rates(Q2, args.n2, args.steps, args.warmup)
        del Q2
        torch.cuda.empty_cache()
        box2 = {"workload": f"hydro step, {args.n2}^3 turbulent box (sigma_ln_rho={SIGMA2}, rms Mach {MACH2})",
                "mixed": {"value": round(r2["mixed"]["mcells"], 2), "ms_per_step": round(r2["mixed"]["ms_per_step"], 4),
                          "stage1_ms": round(r2["mixed"]["stage1_ms"], 4), "stage2_ms": round(r2["mixed"]["stage2_ms"], 4)},
                "fp64": {"value": round(r2["fp64"]["mcells"], 2), "ms_per_step": round(r2["fp64"]["ms_per_step"], 4)},
                "mixed_over_fp64": round(r2["mixed"]["mcells"] / r2["fp64"]["mcells"], 3)}
    cool = None
    if not args.no_cooling:
        t0 = time.perf_counter()
        cool = cooling_rates(args.cooling_log2, args.surrogate, args.gate)
        log(f"cooling {cool} ({time.perf_counter() - t0:.1f}s)")
    cl = None
    if args.cloud_n > 0:
        t0 = time.perf_counter()
        cl = cloud_rates(args.cloud_n, args.cloud_steps, args.cloud_surrogate, args.gate)
        log(f"cloud {cl} ({time.perf_counter() - t0:.1f}s)")
    star = None
    if not args.no_star:
        from torch_amd import star_bench

        t0 = time.perf_counter()
        star = star_bench.gpu_rates(args.star_n, 5)
        if not args.no_cpu:
            star["cpu_baseline"] = star_cpu_baseline(48)
        log(f"star {star} ({time.perf_counter() - t0:.1f}s)")
    ncell = n ** 3
    achieved = BYTES_STAGE2 * ncell / (mx["stage2_ms"] * 1e-3) / 1e9
    # the kernel instance with 32-bit byte offsets runs on grids of fewer than 2^29 cells
    a32 = "true" if ncell * 8 < (1 << 32) else "false"
    traffic = measured_traffic(n, f"void tg::hydro_stage_mixed_kernel<2, true, 2, {a32}>")
    out = {
        "metric": "Mcell-updates/s: FP64 vs mixed hydro; surrogate vs subcycled cooling",
        "value": round(mx["mcells"], 2),
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(mx["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 flux / f64 update",
        "data": "synthetic: seeded log-normal turbulent field (PCG64 seeds 1001-1004, P(k)~k^-11/3, 2<=|k|<=32)",
        "config": {"workload": f"hydro step, {n}^3 turbulent box (sigma_ln_rho={SIGMA}, rms Mach {MACH}), "
                               "RotatedHLLC + albada, outflow", "grid": n, "gamma": GAMMA,
                   "temperature_floor": TFLOOR, "debug": False},
        "fp64": {"value": round(f64["mcells"], 2), "ms_per_step": round(f64["ms_per_step"], 4),
                 "stage1_ms": round(f64["stage1_ms"], 4), "stage2_ms": round(f64["stage2_ms"], 4)},
        "mixed": {"stage1_ms": round(mx["stage1_ms"], 4), "stage2_ms": round(mx["stage2_ms"], 4)},
        "mixed_over_fp64": round(mx["mcells"] / f64["mcells"], 3),
        "mixed_vs_fp64_error": errs,
        "roofline": {"bound": "hbm", "kernel": "hydro_stage_mixed_kernel<rotated,albada,corrector>",
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "bytes_per_cell": BYTES_STAGE2, "launch_ms": round(mx["stage2_ms"], 4)},
        "cpu_baseline": cpu,
        "box256": box2,
        "cooling": cool,
        "cloud": cl,
        "star": star,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
