"""Operator-split hydro + cooling on the cloud set-up (the 384^3 workload).

    python -m torch_amd.cloud [--n 384] [--steps 100] [--precision mixed]
                              [--cooling surrogate|fp64] [--surrogate PATH] [--host-loop]

Set-up (cgs, converted with the scales 1e-20 g/cm^3, 1e-28 dyn/cm^2, 1e11 s):
a cube of side 0.5 pc with an ionising source at its centre; hydrogen number
density n(R) = n0 / (1 + R^2 / rc^2) about a cloud centre 0.35 pc from the
source along +x, rc = 0.01 pc, n0 such that n = 3.2e4 cm^-3 at R = 0.35 pc
(i.e. at the source); T = 300 K; v = 0; HII fraction 1 within 0.1 pc of the
source, else 0; ADV = 1 (the thermo_hii_switch gate passes).  The ionising
flux is F = S exp(-tau) / (4 pi r^2) with S = 4.9e48 s^-1 and tau = sigma
N(r), sigma = 6.3e-18 cm^2, N(r) the analytic hydrogen column from the
source along the ray: with u = r_hat . c and b^2 = rc^2 + |c|^2 - u^2,
N = n0 rc^2 / b (atan((r - u) / b) + atan(u / b)).  The cooling operator
takes F_FUV = 0.5 F / 1.2e7 (the reference's fluxFUV(0.5 S, r^2) with the
attenuated flux) and COL_DEN = N (A_V = 1.086 sigma_dust N).  The source's
cell is excluded from cooling as the reference's CausalWind star cell.

Step: Torch::fullStep with hydro and thermodynamics (Torch.cpp:389-418):
dt = min(hydro CFL, thermo dt) (first step dt_max 1e-20), then the Strang
sequence alternating T(dt/2) H(dt) T(dt/2) and H(dt/2) T(dt) H(dt/2) with
single-stage hydro sub-steps; the thermo sub-steps use the surrogate (FP64
prepass, BF16 MLP, FP64 fallback gated at 1% estimated error) or the FP64
subcycled integrator.  By default the whole loop runs on the device
(torchgpu_run_split: dt reductions and the sequence enqueued without host
round trips); --host-loop drives each step from the host (HostLoop), which
reads dt back as the reference's driver does.
"""
from __future__ import annotations

import argparse
import json
import os
import time

import numpy as np

PC = 3.0857e18
M_H = 1.674e-24
R_CGS = 8.314462e7
SCALES = (1.0e-20, 1.0e-28, 1.0e11)


def setup(n: int, to_code):
    """(Q [7][n][n][n], col_den [n^3], F_FUV [n^3], excluded flat index,
    side length), all in code units."""
    side = 0.5 * PC
    dx = side / n
    xs = (np.arange(n) + 0.5) * dx - 0.5 * side  # cell centres relative to the source
    z, y, x = np.meshgrid(xs, xs, xs, indexing="ij")
    r2 = x * x + y * y + z * z
    r = np.sqrt(r2)
    rc = 0.01 * PC
    cx = 0.35 * PC
    n0 = 3.2e4 * (1.0 + (cx / rc) ** 2)
    R2 = (x - cx) ** 2 + y * y + z * z
    nH = n0 / (1.0 + R2 / (rc * rc))
    u = x / r * cx                                    # r_hat . c
    b = np.sqrt(rc * rc + cx * cx - u * u)
    N = n0 * rc * rc / b * (np.arctan((r - u) / b) + np.arctan(u / b))
    F = 4.9e48 * np.exp(-6.3e-18 * N) / (4.0 * np.pi * r2)
    hii = (r < 0.1 * PC).astype(np.float64)
    rho = nH * M_H
    pre = rho * R_CGS * (hii + 1.0) * 300.0          # X = 1: mu_inv = x + 1
    Q = np.zeros((7, n, n, n))
    Q[0] = rho * to_code(1.0, 1, -3, 0)
    Q[1] = pre * to_code(1.0, 1, -1, -2)
    Q[2] = hii
    Q[3] = 1.0
    col = (N * to_code(1.0, 0, -2, 0)).ravel()
    ffuv = (0.5 * F / 1.2e7 * to_code(1.0, 0, -2, -1)).ravel()
    c = n // 2
    excluded = c + n * (c + n * c)
    return Q, col, ffuv, excluded, side * to_code(1.0, 0, 1, 0)


def cooling_inputs(g, p, col_den, F_FUV, excluded, dt):
    """[9][n] device inputs of the cooling step for the grid's current state,
    formed as the library's grid gather forms them (Q = fixPrimitives(QfromU(U))
    with the HII/ADV clamps and floors, E = U[PRE], T_min 100 K default, the
    excluded cell's ADV = -1)."""
    import torch

    n = p.nx * p.ny * p.nz
    U = torch.empty((7, n), dtype=torch.float64, device="cuda")
    g.get_conserved(out=U)
    den = U[0]
    ke = 0.5 * (U[4:7] * U[4:7]).sum(0) / den
    pre = ((U[1] - ke) * (p.gamma - 1.0)).clamp_min(p.pfloor)
    hii = (U[2] / den).clamp(0.0, 1.0)
    adv = (U[3] / den).clamp(0.0, 1.0)
    adv[excluded] = -1.0
    dc = den.clamp_min(p.dfloor)
    mu_inv = p.mass_fraction_h * (hii + 1.0) + (1.0 - p.mass_fraction_h) * 0.25
    T = pre / (mu_inv * p.spec_gas_const * dc)
    pre = torch.where(T < p.tfloor, mu_inv * p.spec_gas_const * den * p.tfloor, pre)
    tmin = torch.full_like(den, 100.0)
    return torch.stack([dc, pre, hii, adv, U[1], tmin, col_den, F_FUV, torch.full_like(den, dt)]).contiguous()


def build(n: int, precision: str = "mixed", tmax_s: float = 1.0e13):
    """The cloud set-up on the GPU: (params, Grid, Cooling, thermo keyword
    arguments {col_den, F_FUV, excluded} as device tensors)."""
    import torch

    import torch_amd as ta

    tc = lambda v, m, l, t: ta.units(SCALES, v, m, l, t)  # noqa: E731
    Q, col, ffuv, excluded, side = setup(n, tc)
    dfloor = tc(1.0e-30, 1, -3, 0)
    R = tc(R_CGS, 0, 2, -2)
    tfloor = 0.1
    p = ta.make_params((n, n, n), side_length=side, gamma=1.67, dfloor=dfloor, pfloor=0.1 * R * dfloor * tfloor,
                       tfloor=tfloor, spec_gas_const=R, tmax=tc(tmax_s, 0, 0, 1),
                       precision=ta.MIXED if precision == "mixed" else ta.FP64)
    g = ta.Grid(p)
    g.set_primitive(np.ascontiguousarray(Q))
    cool = ta.Cooling(ta.thermo_params(scales=SCALES))
    thermo = dict(col_den=torch.from_numpy(col).cuda(), F_FUV=torch.from_numpy(ffuv).cuda(), excluded=excluded)
    return p, g, cool, thermo


class HostLoop:
    """Torch::fullStep driven from the host: dt from the two operators' time
    steps (host round trips), the Strang sequence as separate calls."""

    def __init__(self, g, cool, thermo, dt_max, surrogate):
        self.g, self.cool, self.thermo, self.dt_max, self.surrogate = g, cool, thermo, dt_max, surrogate
        self.first = True
        self.counter = 0

    def step(self) -> float:
        # Torch::calculateTimeStep
        if self.first:
            dt = self.dt_max * 1.0e-20
            self.first = False
        else:
            dt = min(self.g.hydro_timestep(self.dt_max), self.cool.grid_timestep(self.g, self.dt_max, **self.thermo))
        # Strang alternation of Torch::fullStep over (hydro, thermo)
        self.counter = (self.counter + 1) % 2
        comps = ["hydro", "thermo"]
        seq = [(comps[(0 + self.counter) % 2], 0.5 * dt), (comps[(1 + self.counter) % 2], dt),
               (comps[(0 + self.counter) % 2], dt / 2.0)]
        for comp, h in seq:
            if comp == "hydro":
                self.g.stage_dt(0, h)
            else:
                self.cool.grid_substep(self.g, h, surrogate=self.surrogate, **self.thermo)
        return dt


def main():
    import torch

    import torch_amd as ta
    from torch_amd import surrogate as sg

    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--n", type=int, default=384)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--precision", choices=["mixed", "fp64"], default="mixed")
    ap.add_argument("--cooling", choices=["surrogate", "fp64"], default="surrogate")
    ap.add_argument("--surrogate", default=os.path.join(os.path.dirname(os.path.abspath(__file__)), "data",
                                                        "surrogate_cloud.npz"),
                    help="network weights (default: the one trained on the cloud's conditions)")
    ap.add_argument("--gate", type=float, default=0.01)
    ap.add_argument("--tmax", type=float, default=1.0e13, help="simulation time, s (dt_max = tmax / 10)")
    ap.add_argument("--every", type=int, default=10, help="report (and check for step errors) every N steps")
    ap.add_argument("--check", default="", help="comma-separated steps after which the surrogate and the FP64 "
                    "integrator are compared on the grid's cells (untimed)")
    ap.add_argument("--host-loop", action="store_true",
                    help="drive each step from the host (dt read back) instead of torchgpu_run_split")
    a = ap.parse_args()
    n = a.n
    t0 = time.perf_counter()
    p, g, cool, thermo = build(n, a.precision, a.tmax)
    mlp = None
    use_sur = a.cooling == "surrogate"
    if use_sur:
        mlp = sg.attach(cool, a.surrogate, gate=a.gate)
    torch.cuda.synchronize()
    print(json.dumps({"setup": f"cloud {n}^3", "seconds": round(time.perf_counter() - t0, 1)}), flush=True)
    loop = HostLoop(g, cool, thermo, p.dt_max, use_sur)
    checks = {int(x) for x in a.check.split(",") if x}
    stops = sorted({s for s in range(a.every, a.steps + 1, a.every)} | {a.steps} | {c for c in checks if c <= a.steps})
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ms = 0.0
    t_wall = time.perf_counter()
    done = 0
    sim_t = 0.0
    for stop in stops:
        ev0.record()
        if a.host_loop:
            for _ in range(stop - done):
                last_dt = loop.step()
                sim_t += last_dt
        else:
            cool.run_split(g, stop - done, surrogate=use_sur, **thermo)
        ev1.record()
        g.synchronize()
        ms += ev0.elapsed_time(ev1)
        done = stop
        t_dev, dt, nst = g.clock()
        if a.host_loop:
            t_dev, dt = sim_t, last_dt
        print(json.dumps({"step": stop, "dt": dt, "time": t_dev,
                          "elapsed_s": round(time.perf_counter() - t_wall, 2)}), flush=True)
        if stop in checks and use_sur:
            inp = cooling_inputs(g, p, thermo["col_den"], thermo["F_FUV"], thermo["excluded"], dt)
            r = sg.compare(cool, inp)
            r["out_of_envelope"] = r["cells"] - r["exact"] - r["mlp_queued"]
            print(json.dumps({"step": stop, "cooling_check": r}), flush=True)
            del inp
            torch.cuda.empty_cache()
    print(json.dumps({"workload": f"cloud {n}^3, {a.steps} operator-split steps, {a.precision} hydro, "
                                  f"{a.cooling} cooling, {'host' if a.host_loop else 'device'} loop",
                      "ms_per_step": round(ms / a.steps, 3),
                      "Mcell_updates_per_s": round(n ** 3 * a.steps / (ms * 1e-3) / 1e6, 2)}), flush=True)
    del mlp


if __name__ == "__main__":
    main()
