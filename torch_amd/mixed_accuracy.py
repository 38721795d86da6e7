"""Mixed-precision hydro against the FP64 GPU path on the turbulent box.

    python -m torch_amd.mixed_accuracy [--n 512] [--steps 200] [--sigma 1.5] [--mach 4]

Both grids start from the same seeded log-normal turbulent field
(torch_amd.setups.turbulent), outflow boundaries, RotatedHLLC + van Albada,
2nd order in space and time.  Both take the same dt every step: the first
step dt_max * 1e-20, then the smaller of the two grids' CFL 0.5 limits
(Hydrodynamics::calculateTimeStep on each state), so the two differ only by
the precision of the reconstruction and the fluxes and neither exceeds its
own CFL limit (in the strong rarefactions of the Mach 4 field the two
states' largest signal speeds differ).  Printed: relative L1 and Linf errors
of density, pressure and velocity (|v| components pooled) every `--every`
steps and at the end, one JSON object per line; a step error of either run
(the reference's exceptions) is reported and ends the run.
"""
from __future__ import annotations

import argparse
import json
import time

R_CGS = 8.314462e7


def errors(q64, qmx):
    out = {}
    for key, idx in (("density", [0]), ("pressure", [1]), ("velocity", [4, 5, 6])):
        a, b = q64[idx], qmx[idx]
        d = (b - a).abs()
        out[key] = {"rel_l1": float(d.sum() / a.abs().sum()), "rel_linf": float(d.max() / a.abs().max())}
    return out


def main():
    import torch

    import torch_amd as ta
    from torch_amd import setups

    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--sigma", type=float, default=1.5)
    ap.add_argument("--mach", type=float, default=4.0)
    ap.add_argument("--gamma", type=float, default=1.67)
    ap.add_argument("--tfloor", type=float, default=1e-12)
    ap.add_argument("--every", type=int, default=50)
    a = ap.parse_args()
    n = a.n
    pfloor = 0.1 * R_CGS * 1e-30 * a.tfloor
    t0 = time.perf_counter()
    Q = torch.from_numpy(setups.turbulent(n, a.sigma, a.mach, gamma=a.gamma, device="cuda")).cuda()
    print(json.dumps({"setup": f"{n}^3 turbulent box sigma_ln_rho={a.sigma} rms Mach {a.mach}",
                      "ic_seconds": round(time.perf_counter() - t0, 1)}), flush=True)
    # debug = false as the reference's shipped configuration (config/torch-config.lua)
    kw = dict(gamma=a.gamma, tfloor=a.tfloor, pfloor=pfloor, debug=False)
    g64 = ta.Grid(ta.make_params((n, n, n), precision=ta.FP64, **kw))
    gmx = ta.Grid(ta.make_params((n, n, n), precision=ta.MIXED, **kw))
    g64.set_primitive(Q)
    gmx.set_primitive(Q)
    del Q
    torch.cuda.empty_cache()
    q64 = torch.empty((7, n, n, n), dtype=torch.float64, device="cuda")
    qmx = torch.empty_like(q64)
    t0 = time.perf_counter()
    dt_max = g64.params.dt_max
    sim_t = 0.0
    for s in range(1, a.steps + 1):
        if s == 1:
            dt = dt_max * 1.0e-20
        else:
            dt = min(g64.hydro_timestep(dt_max), gmx.hydro_timestep(dt_max))
        g64.hydro_step(dt)
        gmx.hydro_step(dt)
        sim_t += dt
        if s % a.every == 0 or s == a.steps:
            for name, g in (("fp64", g64), ("mixed", gmx)):
                try:
                    g.synchronize()
                except ta.TorchGpuError as e:
                    print(json.dumps({"step": s, "run": name, "error": str(e)}), flush=True)
                    return
            g64.get_primitive(fix=False, out=q64)
            gmx.get_primitive(fix=False, out=qmx)
            print(json.dumps({"step": s, "time": sim_t, "dt": dt, "errors": errors(q64, qmx),
                              "elapsed_s": round(time.perf_counter() - t0, 1)}), flush=True)
    g64.close()
    gmx.close()


if __name__ == "__main__":
    main()
