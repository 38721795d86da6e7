"""Torch::fullStep driven from the host with any of the reference's active
components: hydrodynamics, thermodynamics (cooling) and radiation
(src/Torch/Torch.cpp:240-418).

    loop = FullStep(grid, cooling=cool, radiation=rad)
    for _ in range(nsteps):
        dt = loop.step()

Per step, as Torch::run + fullStep + calculateTimeStep + subStep:
  - the checkpoint clip dt_nextCheckpoint (Checkpointer::update,
    src/IO/Checkpointer.cpp:14-26) at the current time;
  - preTimeStepCalculations of thermodynamics and radiation on the state
    (the library's thermo calls gather their inputs themselves);
  - dt = dt_max * 1e-20 on the first step, else min(min(dt_hydro, dt_rad),
    dt_thermo), then min(dt_nextCheckpoint, dt);
  - with one component hydroStep(dt) (the library's two-stage step);
    otherwise stepCounter = (stepCounter + 1) % ncomps and the sub-steps
    h dt (h = 1/2, ..., 1 for the last) over the components rotated by
    stepCounter, then dt / 2.0 back down (Torch.cpp:406-415), each a
    single-stage sub-step; the first reuses the pre-step heating.
Components are ordered hydro, thermo, radiation (activeComponents,
Torch.cpp:256-261).
"""
from __future__ import annotations


class FullStep:
    def __init__(self, grid, cooling=None, thermo: dict | None = None, radiation=None, surrogate: bool = False):
        p = grid.params
        self.g, self.cool, self.rad = grid, cooling, radiation
        self.thermo = dict(thermo or {})
        self.surrogate = surrogate
        self.dt_max, self.tmax, self.ncheckpoints = p.dt_max, p.tmax, p.ncheckpoints
        self.first = True
        self.counter = 0
        self.time = 0.0
        self.count = 0
        self.comps = ["hydro"] + (["thermo"] if cooling is not None else []) + (["rad"] if radiation else [])

    def _checkpoint_dt(self) -> float:
        delta = 100.0 / self.ncheckpoints
        pct = min(100.0, 100.0 * self.time / self.tmax)
        while pct - self.count * delta > delta - 1.0e-8:
            self.count += 1
        return ((self.count + 1) * delta - pct) * self.tmax / 100.0

    def _sub(self, comp: str, h: float, has_heat: bool):
        if comp == "hydro":
            self.g.stage_dt(0, h)
        elif comp == "thermo":
            self.cool.grid_substep(self.g, h, surrogate=self.surrogate, **self.thermo)
        else:
            if not has_heat:
                self.rad.pre_step()
            self.rad.substep(h, True)

    def step(self) -> float:
        dt_next = self._checkpoint_dt()
        if self.rad is not None:
            self.rad.pre_step()
        if self.first:
            dt = self.dt_max * 1.0e-20
            self.first = False
        else:
            dt_hydro = self.g.hydro_timestep(self.dt_max)
            dt_rad = self.rad.timestep(self.dt_max) if self.rad is not None else dt_hydro
            dt_thermo = (self.cool.grid_timestep(self.g, self.dt_max, **self.thermo) if self.cool is not None
                         else dt_hydro)
            dt = min(min(dt_hydro, dt_rad), dt_thermo)
        dt = min(dt_next, dt)
        n = len(self.comps)
        if n == 1:
            self.g.hydro_step(dt)
        else:
            self.counter = (self.counter + 1) % n
            for i in range(n):
                self._sub(self.comps[(i + self.counter) % n], 1.0 * dt if i == n - 1 else 0.5 * dt, i == 0)
            for i in range(n - 2, -1, -1):
                self._sub(self.comps[(i + self.counter) % n], dt / 2.0, False)
        self.time += dt
        return dt
