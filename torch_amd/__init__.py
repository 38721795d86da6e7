"""torch_amd -- MI355X (gfx950) hydro + thermodynamics step of TORCH.

Python binding of the C ABI in include/torchgpu.h (libtorchgpu.so, built
in-tree by torch_amd/build.py).  There is no CPU fallback: if the HIP
library is missing, importing the binding raises.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtorchgpu.so")

OK = 0
ERRORS = {
    1: "invalid argument", 2: "HIP error", 3: "RiemannSolver::solve: HII flux is NaN",
    4: "Fluid::fixSolution: invalid value", 5: "hydro before produced an error",
    6: "Hydrodynamics::calcFluxes: invalid order", 7: "RiemannSolverFactory::create: unknown type",
    8: "SlopeLimiterFactory::create: unknown type", 9: "subcycle count out of range",
    10: "surrogate not trained", 11: "radiation error",
}
BC = {"free": 0, "reflecting": 1, "outflow": 2, "inflow": 3, "periodic": 4, "default": 0}
SOLVER = {"HLL": 0, "HLLC": 1, "RotatedHLLC": 2, "default": 2}
LIMITER = {"monotonised_central": 0, "superbee": 1, "minmod": 2, "maxmod": 3,
           "leer": 4, "ospre": 5, "albada": 6, "default": 6}
FP64, MIXED = 0, 1


class TorchGpuError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class Params(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
        ("spatial_order", ctypes.c_int),
        ("bc_left", ctypes.c_int * 3), ("bc_right", ctypes.c_int * 3),
        ("riemann_solver", ctypes.c_int), ("slope_limiter", ctypes.c_int),
        ("precision", ctypes.c_int), ("ncheckpoints", ctypes.c_int), ("debug", ctypes.c_int),
        ("side_length", ctypes.c_double), ("gamma", ctypes.c_double),
        ("dfloor", ctypes.c_double), ("pfloor", ctypes.c_double), ("tfloor", ctypes.c_double),
        ("mass_fraction_h", ctypes.c_double), ("spec_gas_const", ctypes.c_double),
        ("tmax", ctypes.c_double), ("dt_max", ctypes.c_double),
    ]


_lib = None


def lib() -> ctypes.CDLL:
    """The loaded HIP library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run `python -m torch_amd.build` (there is no CPU fallback)")
        L = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        d = ctypes.c_double
        i = ctypes.c_int
        L.torchgpu_create.argtypes = [ctypes.POINTER(Params), vp, ctypes.POINTER(vp)]
        L.torchgpu_destroy.argtypes = [vp]
        L.torchgpu_destroy.restype = None
        L.torchgpu_last_error.argtypes = [vp, ctypes.c_char_p, ctypes.c_size_t]
        L.torchgpu_state_ptr.argtypes = [vp, ctypes.POINTER(vp)]
        L.torchgpu_set_time.argtypes = [vp, d]
        L.torchgpu_pause_at_checkpoints.argtypes = [vp, i]
        L.torchgpu_checkpoint.argtypes = [vp, ctypes.POINTER(i), ctypes.POINTER(i)]
        L.torchgpu_get_print_state.argtypes = [vp, vp, i]
        for name in ("torchgpu_set_primitive", "torchgpu_set_conserved", "torchgpu_get_conserved"):
            getattr(L, name).argtypes = [vp, vp, i]
        L.torchgpu_get_primitive.argtypes = [vp, vp, i, i]
        L.torchgpu_hydro_timestep.argtypes = [vp, d, ctypes.POINTER(d)]
        L.torchgpu_hydro_step.argtypes = [vp, d]
        L.torchgpu_run.argtypes = [vp, i]
        L.torchgpu_hydro_stage.argtypes = [vp, i]
        L.torchgpu_hydro_stage_dt.argtypes = [vp, i, d]
        L.torchgpu_set_dt.argtypes = [vp, d]
        L.torchgpu_reset_clock.argtypes = [vp]
        L.torchgpu_synchronize.argtypes = [vp]
        L.torchgpu_get_clock.argtypes = [vp, ctypes.POINTER(d), ctypes.POINTER(d), ctypes.POINTER(i)]
        L.torchgpu_is_quitting.argtypes = [vp, ctypes.POINTER(i)]
        L.torchgpu_heating_track.argtypes = [vp, i]
        _lib = L
    return _lib


def _current_device() -> int:
    import torch

    return torch.cuda.current_device()


def default_stream(stream: int | None = None) -> int:
    """The HIP stream a handle enqueues on: the one given, else torch's
    current stream on the current device, so that library work is ordered
    with the torch tensors it reads and writes (the library's own stream
    would race with them)."""
    if stream is not None:
        return stream
    import torch

    return torch.cuda.current_stream().cuda_stream


def make_params(shape, side_length=1.0, gamma=1.67, dfloor=1e-30, pfloor=0.0, tfloor=0.0,
                mass_fraction_h=1.0, spec_gas_const=8.314462e7, tmax=1e3, dt_max=None,
                ncheckpoints=1, solver="RotatedHLLC", limiter="albada", spatial_order=1,
                bc="outflow", precision=FP64, debug=True) -> Params:
    """Parameters in code units.  shape = (nz, ny, nx).  `debug` is
    Integration.debug, true by default as in TorchParameters
    (Parameters.hpp:50): the time loop stops after a step whose hydro dt is
    at most 1e-8 tmax (Torch.cpp:321-332)."""
    p = Params()
    p.nz, p.ny, p.nx = shape
    p.spatial_order = spatial_order
    bcs = bc if isinstance(bc, (list, tuple)) else [bc] * 6
    for k in range(3):
        p.bc_left[k] = BC[bcs[2 * k]]
        p.bc_right[k] = BC[bcs[2 * k + 1]]
    p.riemann_solver = SOLVER[solver] if isinstance(solver, str) else solver
    p.slope_limiter = LIMITER[limiter] if isinstance(limiter, str) else limiter
    p.precision = precision
    p.ncheckpoints = ncheckpoints
    p.debug = 1 if debug else 0
    p.side_length, p.gamma = side_length, gamma
    p.dfloor, p.pfloor, p.tfloor = dfloor, pfloor, tfloor
    p.mass_fraction_h, p.spec_gas_const = mass_fraction_h, spec_gas_const
    p.tmax = tmax
    p.dt_max = tmax / 10.0 if dt_max is None else dt_max
    return p


def _ptr(a) -> int:
    """Address of a contiguous float64 numpy array or torch tensor."""
    if isinstance(a, np.ndarray):
        assert a.dtype == np.float64 and a.flags.c_contiguous
        return a.ctypes.data
    assert a.dtype.__str__() == "torch.float64" and a.is_contiguous()
    return a.data_ptr()


class Grid:
    """One whole grid on the current GPU (a torchgpu_ctx)."""

    def __init__(self, params: Params, stream: int | None = None):
        self.params = params
        self.shape = (params.nz, params.ny, params.nx)
        self._h = ctypes.c_void_p()
        rc = lib().torchgpu_create(ctypes.byref(params), ctypes.c_void_p(default_stream(stream)),
                                   ctypes.byref(self._h))
        if rc:
            raise TorchGpuError(rc, ERRORS.get(rc, "create failed"))

    def _check(self, rc: int):
        if rc:
            buf = ctypes.create_string_buffer(4096)
            lib().torchgpu_last_error(self._h, buf, 4096)
            raise TorchGpuError(rc, buf.value.decode() or ERRORS.get(rc, "error"))

    def close(self):
        if self._h:
            lib().torchgpu_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- state -------------------------------------------------------------
    def _is_dev(self, a) -> int:
        return 0 if isinstance(a, np.ndarray) else int(a.is_cuda)

    def set_primitive(self, Q):
        self._check(lib().torchgpu_set_primitive(self._h, ctypes.c_void_p(_ptr(Q)), self._is_dev(Q)))

    def set_conserved(self, U):
        self._check(lib().torchgpu_set_conserved(self._h, ctypes.c_void_p(_ptr(U)), self._is_dev(U)))

    def get_conserved(self, out=None):
        out = np.empty((7, *self.shape)) if out is None else out
        self._check(lib().torchgpu_get_conserved(self._h, ctypes.c_void_p(_ptr(out)), self._is_dev(out)))
        return out

    def get_primitive(self, fix: bool = True, out=None):
        out = np.empty((7, *self.shape)) if out is None else out
        self._check(lib().torchgpu_get_primitive(self._h, ctypes.c_void_p(_ptr(out)), int(fix), self._is_dev(out)))
        return out

    def state_ptr(self) -> int:
        p = ctypes.c_void_p()
        self._check(lib().torchgpu_state_ptr(self._h, ctypes.byref(p)))
        return p.value

    # -- operators ---------------------------------------------------------
    def hydro_timestep(self, dt_max: float | None = None) -> float:
        dt = ctypes.c_double()
        self._check(lib().torchgpu_hydro_timestep(self._h, self.params.dt_max if dt_max is None else dt_max,
                                                  ctypes.byref(dt)))
        return dt.value

    def hydro_step(self, dt: float):
        self._check(lib().torchgpu_hydro_step(self._h, dt))

    def run(self, nsteps: int):
        self._check(lib().torchgpu_run(self._h, nsteps))

    def set_dt(self, dt: float):
        self._check(lib().torchgpu_set_dt(self._h, dt))

    def stage(self, stage: int):
        self._check(lib().torchgpu_hydro_stage(self._h, stage))

    def stage_dt(self, stage: int, dtf: float):
        """One stage with an explicit, already scaled dt; stage 0 is the
        single-stage hydro sub-step of an operator-split step."""
        self._check(lib().torchgpu_hydro_stage_dt(self._h, stage, dtf))

    def reset_clock(self):
        self._check(lib().torchgpu_reset_clock(self._h))

    def set_time(self, t: float):
        """Simulation time (code units) and its checkpoint count, as
        Torch::initialise sets them from a snapshot."""
        self._check(lib().torchgpu_set_time(self._h, t))

    def pause_at_checkpoints(self, on: bool = True):
        self._check(lib().torchgpu_pause_at_checkpoints(self._h, int(on)))

    def checkpoint(self):
        """(checkpoint count, due): due when a run paused at a new
        checkpoint (the pause is lifted by this call)."""
        count, due = ctypes.c_int(), ctypes.c_int()
        self._check(lib().torchgpu_checkpoint(self._h, ctypes.byref(count), ctypes.byref(due)))
        return count.value, bool(due.value)

    def print_state(self, out=None):
        """The primitive state Torch::run prints at this point (see
        torchgpu_get_print_state)."""
        out = np.empty((7, *self.shape)) if out is None else out
        self._check(lib().torchgpu_get_print_state(self._h, ctypes.c_void_p(_ptr(out)), self._is_dev(out)))
        return out

    def synchronize(self):
        self._check(lib().torchgpu_synchronize(self._h))

    def clock(self):
        t, dt, n = ctypes.c_double(), ctypes.c_double(), ctypes.c_int()
        self._check(lib().torchgpu_get_clock(self._h, ctypes.byref(t), ctypes.byref(dt), ctypes.byref(n)))
        return t.value, dt.value, n.value

    def heating_track(self, on: bool = True):
        """Keep the GridCell::H terms that persist between heating fills
        (torchgpu_heating_track)."""
        self._check(lib().torchgpu_heating_track(self._h, int(on)))

    def quitting(self) -> bool:
        """Torch::isQuitting: the debug test of Torch::calculateTimeStep
        ended the time loop (Torch.cpp:321-332)."""
        q = ctypes.c_int()
        self._check(lib().torchgpu_is_quitting(self._h, ctypes.byref(q)))
        return bool(q.value)


# -- cooling operator -------------------------------------------------------
class ThermoParams(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in (
        "thermo_hii_switch", "heating_amplification", "mass_fraction_h", "gamma", "pfloor",
        "hydrogen_mass", "spec_gas_const", "dust_xsec", "pi",
        "z0", "T1", "T2", "T3", "T4", "imlc", "nmlc", "ciec_minT", "ciec", "cxhi_damp", "n0", "nmc",
        "fuvh_a", "fuvh_b", "fuvh_c", "irh_a", "irh_b", "crh", "spl_min_slope", "spl_max_slope")] + [
        ("spl_x", ctypes.c_double * 26), ("spl_y", ctypes.c_double * 26), ("spl_y2", ctypes.c_double * 26),
        ("subcycling", ctypes.c_int), ("pad", ctypes.c_int)]


class Star(ctypes.Structure):
    """torchgpu_star: Parameters.Star cell position, wind radius (cells) and
    photon rate (code units)."""
    _fields_ = [("cell", ctypes.c_int * 3), ("wind_cell_radius", ctypes.c_int), ("photon_rate", ctypes.c_double)]


def make_star(cell, wind_cell_radius: int = 0, photon_rate: float = 0.0) -> Star:
    """cell = (i, j, k), x first."""
    s = Star()
    for a in range(3):
        s.cell[a] = int(cell[a])
    s.wind_cell_radius = int(wind_cell_radius)
    s.photon_rate = float(photon_rate)
    return s


def _bind_cooling(L):
    vp, d, i = ctypes.c_void_p, ctypes.c_double, ctypes.c_int
    L.torchgpu_units.argtypes = [d, d, d, d, d, d, d, i, ctypes.POINTER(d)]
    L.torchgpu_thermo_params_init.argtypes = [ctypes.POINTER(ThermoParams), d, d, d, d, d, d, d, d, d, i]
    L.torchgpu_cooling_create.argtypes = [ctypes.POINTER(ThermoParams), vp, ctypes.POINTER(vp)]
    L.torchgpu_cooling_destroy.argtypes = [vp]
    L.torchgpu_cooling_destroy.restype = None
    L.torchgpu_cooling_last_error.argtypes = [vp, ctypes.c_char_p, ctypes.c_size_t]
    L.torchgpu_thermo_cells.argtypes = [vp, ctypes.c_int64, vp, vp, ctypes.POINTER(ctypes.c_int64)]
    L.torchgpu_thermo_grid_timestep.argtypes = [vp, vp, vp, vp, vp, ctypes.c_int64, d, ctypes.POINTER(d)]
    L.torchgpu_thermo_grid_substep.argtypes = [vp, vp, vp, vp, vp, ctypes.c_int64, d, i]
    L.torchgpu_run_split.argtypes = [vp, vp, vp, vp, vp, ctypes.c_int64, i, i]
    L.torchgpu_cooling_set_star.argtypes = [vp, vp, ctypes.POINTER(Star)]
    L.torchgpu_cooling_set_min_temp.argtypes = [vp, vp, ctypes.c_int64, i]
    L.torchgpu_star_geometry.argtypes = [vp, vp, vp, vp, vp, i]
    L.torchgpu_thermo_grid_inputs.argtypes = [vp, vp, vp, vp, vp, ctypes.c_int64, vp, vp, i]
    L.torchgpu_fill_heating.argtypes = [vp, vp, vp, vp, ctypes.c_int64, vp, i]


def units(scales, value, m, l, t, to_code=True) -> float:
    L = lib()
    if not hasattr(L, "_cooling_bound"):
        _bind_cooling(L)
        L._cooling_bound = True
    out = ctypes.c_double()
    rc = L.torchgpu_units(scales[0], scales[1], scales[2], value, m, l, t, int(to_code), ctypes.byref(out))
    if rc:
        raise TorchGpuError(rc, "units")
    return out.value


# -- snapshots (DataPrinter::print2D / DataReader text format) ----------------

def _bind_snapshot(L):
    if getattr(L, "_snapshot_bound", False):
        return
    d, i, vp, cp = ctypes.c_double, ctypes.c_int, ctypes.c_void_p, ctypes.c_char_p
    L.torchgpu_snapshot_write.argtypes = [cp, i, d, vp, i, i, i, d, d, d, d]
    L.torchgpu_snapshot_header.argtypes = [cp] + [ctypes.POINTER(d)] + [ctypes.POINTER(i)] * 4 + [ctypes.POINTER(d)]
    L.torchgpu_snapshot_read.argtypes = [cp, vp, i, i, i]
    L.torchgpu_snapshot_patch.argtypes = [cp, vp, i, i, i, d, i, i, i]
    L.torchgpu_heating_write.argtypes = [cp, i, d, vp, i, i, i, d, d, d]
    L._snapshot_bound = True


def heating_write(path: str, t: float, H, scales=(1.0, 1.0, 1.0), gz=None):
    """DataPrinter::printHeating: time t and the heating arrays H (host
    float64 [11][nz][ny][nx], code units) as text (gzip when `gz`, by default
    when `path` ends in .gz)."""
    import numpy as np

    L = lib()
    _bind_snapshot(L)
    H = np.ascontiguousarray(H, dtype=np.float64)
    nz, ny, nx = H.shape[1:]
    gz = path.endswith(".gz") if gz is None else gz
    rc = L.torchgpu_heating_write(path.encode(), int(gz), t, H.ctypes.data, nx, ny, nz, *scales)
    if rc:
        raise TorchGpuError(rc, f"heating_write {path}")


def snapshot_write(path: str, t: float, Q, side_length: float, scales=(1.0, 1.0, 1.0), gz=None):
    """Writes time t and the primitive state Q (host float64 [7][nz][ny][nx],
    code units) in the reference's data2D text format (gzip when `gz`, by
    default when `path` ends in .gz)."""
    import numpy as np

    L = lib()
    _bind_snapshot(L)
    Q = np.ascontiguousarray(Q, dtype=np.float64)
    nz, ny, nx = Q.shape[1:]
    gz = path.endswith(".gz") if gz is None else gz
    rc = L.torchgpu_snapshot_write(path.encode(), int(gz), t, Q.ctypes.data, nx, ny, nz, side_length, *scales)
    if rc:
        raise TorchGpuError(rc, f"snapshot_write: cannot write {path}")


def snapshot_header(path: str) -> dict:
    """Time (s), shape (nz, ny, nx), nd and side length (cm) of a snapshot."""
    L = lib()
    _bind_snapshot(L)
    t, side = ctypes.c_double(), ctypes.c_double()
    nx, ny, nz, nd = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    rc = L.torchgpu_snapshot_header(path.encode(), ctypes.byref(t), ctypes.byref(nx), ctypes.byref(ny),
                                    ctypes.byref(nz), ctypes.byref(nd), ctypes.byref(side))
    if rc:
        raise TorchGpuError(rc, f"snapshot_header: invalid input file {path}")
    return {"time": t.value, "shape": (nz.value, ny.value, nx.value), "nd": nd.value, "side_length": side.value}


def snapshot_patch(path: str, Q, grid_dx: float, offset=(0, 0, 0)):
    """DataReader::patchGrid on host primitives Q [7][nz][ny][nx] (modified
    in place): the patch file's cells averaged into the grid cells they
    cover, at `offset` cells."""
    import numpy as np

    L = lib()
    _bind_snapshot(L)
    assert isinstance(Q, np.ndarray) and Q.dtype == np.float64 and Q.flags.c_contiguous
    nz, ny, nx = Q.shape[1:]
    rc = L.torchgpu_snapshot_patch(path.encode(), Q.ctypes.data, nx, ny, nz, grid_dx, *[int(o) for o in offset])
    if rc:
        raise TorchGpuError(rc, f"DataReader::patchGrid: cannot patch with {path}")
    return Q


def snapshot_read(path: str, shape=None):
    """(Q, header): density, pressure, HII fraction and velocities of the
    snapshot's cells (cgs, as stored) in a zero-initialised [7][nz][ny][nx]
    array of `shape` (default: the snapshot's own)."""
    import numpy as np

    L = lib()
    _bind_snapshot(L)
    h = snapshot_header(path)
    nz, ny, nx = shape or h["shape"]
    Q = np.zeros((7, nz, ny, nx))
    rc = L.torchgpu_snapshot_read(path.encode(), Q.ctypes.data, nx, ny, nz)
    if rc:
        raise TorchGpuError(rc, f"snapshot_read: invalid input file {path}")
    return Q, h


def run_printing(grid: "Grid", nsteps: int, outdir: str, scales=(1.0, 1.0, 1.0), cooling=None, thermo=None,
                 surrogate: bool = False, radiation=None, heating=None):
    """Torch::run with its snapshots (Torch.cpp:240-302): prints
    data2D_<count>.txt.gz (6-digit count) into `outdir` for the initial
    state, at every checkpoint (the run pauses on the device there) and at
    the end when the loop ends (tmax reached, or the debug stop of
    Torch::calculateTimeStep); at most `nsteps` steps (hydro only, or
    operator-split with `cooling` and its `thermo` keyword arguments and/or
    `radiation`).
    With `heating` (a Cooling handle: its thermo constants, star and the
    `thermo` arrays), the heating arrays are tracked from the start and
    printed before each checkpoint's snapshot as heating_<count>.txt.gz
    (Torch::run's fillHeatingArrays + printHeating, Torch.cpp:264,272-276).
    Returns the list of files written."""
    import os

    p = grid.params
    os.makedirs(outdir, exist_ok=True)
    written = []
    hkw = {k: v for k, v in (thermo or {}).items() if k in ("col_den", "F_FUV", "excluded")}

    def emit(count):
        path = os.path.join(outdir, f"data2D_{count:06d}.txt.gz")
        t, _, _ = grid.clock()
        snapshot_write(path, t, grid.print_state(), p.side_length, scales)
        written.append(path)

    def emit_heating(count):
        path = os.path.join(outdir, f"heating_{count:06d}.txt.gz")
        t, _, _ = grid.clock()
        heating_write(path, t, heating.fill_heating(grid, **hkw), scales)
        written.append(path)

    if heating is not None:
        grid.heating_track(True)
    count, _ = grid.checkpoint()
    emit(count)
    grid.pause_at_checkpoints(True)
    if heating is not None:
        heating.fill_heating(grid, **hkw)  # the fill before Torch::run's loop (Torch.cpp:264)
    final_print = False
    done = 0
    try:
        while done < nsteps:
            n0 = grid.clock()[2]
            if radiation is not None:
                radiation.run_full(nsteps - done, cooling=cooling, surrogate=surrogate, **(thermo or {}))
            elif cooling is None:
                grid.run(nsteps - done)
            else:
                cooling.run_split(grid, nsteps - done, surrogate=surrogate, **(thermo or {}))
            grid.synchronize()
            done += grid.clock()[2] - n0
            count, due = grid.checkpoint()
            if not due:
                break
            if heating is not None:
                emit_heating(count)
            emit(count)
            final_print = count != p.ncheckpoints
        # Torch::run's last print once its loop has ended: tmax reached or
        # quitting (Torch.cpp:266,295-297)
        if final_print and (not grid.clock()[0] < p.tmax or grid.quitting()):
            emit(p.ncheckpoints)
    finally:
        grid.pause_at_checkpoints(False)
    return written


def thermo_params(scales=(1e-20, 1e-28, 1e11), thermo_hii_switch=1e-2, heating_amplification=1.0,
                  mass_fraction_h=1.0, gamma=1.67, density_floor=1e-30, temperature_floor=0.1,
                  subcycling=True) -> ThermoParams:
    L = lib()
    if not hasattr(L, "_cooling_bound"):
        _bind_cooling(L)
        L._cooling_bound = True
    p = ThermoParams()
    rc = L.torchgpu_thermo_params_init(ctypes.byref(p), scales[0], scales[1], scales[2], thermo_hii_switch,
                                       heating_amplification, mass_fraction_h, gamma, density_floor,
                                       temperature_floor, int(subcycling))
    if rc:
        raise TorchGpuError(rc, "thermo_params_init")
    return p


class Cooling:
    """The cooling operator on the current GPU (torchgpu_cooling)."""

    def __init__(self, params: ThermoParams, stream: int | None = None):
        L = lib()
        if not hasattr(L, "_cooling_bound"):
            _bind_cooling(L)
            L._cooling_bound = True
        self.params = params
        self._h = ctypes.c_void_p()
        rc = L.torchgpu_cooling_create(ctypes.byref(params), ctypes.c_void_p(default_stream(stream)),
                                       ctypes.byref(self._h))
        if rc:
            raise TorchGpuError(rc, ERRORS.get(rc, "cooling create failed"))

    def _check(self, rc):
        if rc:
            buf = ctypes.create_string_buffer(1024)
            lib().torchgpu_cooling_last_error(self._h, buf, 1024)
            raise TorchGpuError(rc, buf.value.decode() or ERRORS.get(rc, "error"))

    def close(self):
        if self._h:
            lib().torchgpu_cooling_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def _dptr(a, grid: "Grid"):
        """Device address of an optional per-cell array of `grid` (T_min,
        COL_DEN, F_FUV): a contiguous float64 CUDA tensor of one value per
        cell, or None.  Anything else is refused here, before a kernel would
        read it out of bounds."""
        if a is None:
            return ctypes.c_void_p(0)
        p = grid.params
        n = p.nx * p.ny * p.nz
        if not hasattr(a, "data_ptr") or not getattr(a, "is_cuda", False):
            raise ValueError("per-cell thermo arrays must be CUDA tensors")
        if str(a.dtype) != "torch.float64" or not a.is_contiguous() or a.numel() != n:
            raise ValueError(f"per-cell thermo arrays must be contiguous float64 with {n} values "
                             f"(got {a.dtype}, {tuple(a.shape)}, contiguous={a.is_contiguous()})")
        if a.device.index != _current_device():
            raise ValueError(f"per-cell thermo array on {a.device}, the grid is on cuda:{_current_device()}")
        return ctypes.c_void_p(a.data_ptr())

    def grid_timestep(self, grid: "Grid", dt_max: float, T_min=None, col_den=None, F_FUV=None,
                      excluded: int = -1) -> float:
        """Thermodynamics::calculateTimeStep on the grid state (device
        arrays [n] or None for T_min, COL_DEN, F_FUV)."""
        dt = ctypes.c_double()
        self._check(lib().torchgpu_thermo_grid_timestep(grid._h, self._h, self._dptr(T_min, grid), self._dptr(col_den, grid),
                                                        self._dptr(F_FUV, grid), excluded, dt_max, ctypes.byref(dt)))
        return dt.value

    def grid_substep(self, grid: "Grid", dt: float, T_min=None, col_den=None, F_FUV=None, excluded: int = -1,
                     surrogate: bool = False):
        """The thermo sub-step of Torch::subStep on the grid state."""
        self._check(lib().torchgpu_thermo_grid_substep(grid._h, self._h, self._dptr(T_min, grid), self._dptr(col_den, grid),
                                                       self._dptr(F_FUV, grid), excluded, dt, int(surrogate)))

    def run_split(self, grid: "Grid", nsteps: int, T_min=None, col_den=None, F_FUV=None, excluded: int = -1,
                  surrogate: bool = False):
        """`nsteps` Torch::fullStep iterations of hydro + thermo with dt and
        the Strang alternation on the device (no host round trip)."""
        self._check(lib().torchgpu_run_split(grid._h, self._h, self._dptr(T_min, grid), self._dptr(col_den, grid),
                                             self._dptr(F_FUV, grid), excluded, nsteps, int(surrogate)))

    def set_min_temp(self, T_min=None):
        """Thermodynamics::initialiseMinTempField: the per-cell T_min (host
        or device array [n]) the grid calls use when given none; None
        restores the constant 100 K."""
        if T_min is None:
            self._check(lib().torchgpu_cooling_set_min_temp(self._h, None, 0, 0))
            return
        if hasattr(T_min, "data_ptr"):
            if str(T_min.dtype) != "torch.float64" or not T_min.is_contiguous():
                raise ValueError(f"T_min must be a contiguous float64 tensor (got {T_min.dtype})")
            if T_min.is_cuda and T_min.device.index != _current_device():
                raise ValueError(f"T_min on {T_min.device}, the handle is on cuda:{_current_device()}")
            self._check(lib().torchgpu_cooling_set_min_temp(self._h, ctypes.c_void_p(T_min.data_ptr()),
                                                            T_min.numel(), int(T_min.is_cuda)))
        else:
            import numpy as np

            a = np.ascontiguousarray(T_min, dtype=np.float64).ravel()
            self._check(lib().torchgpu_cooling_set_min_temp(self._h, a.ctypes.data_as(ctypes.c_void_p), a.size, 0))

    def set_star(self, grid: "Grid", star: "Star | None"):
        """Star-on thermodynamics on `grid` (torchgpu_cooling_set_star): ray
        traced COL_DEN, the star's F_FUV, CausalWind cells skipped; None
        switches it off."""
        self._check(lib().torchgpu_cooling_set_star(self._h, grid._h, None if star is None else ctypes.byref(star)))
        p = grid.params
        self._star_n = None if star is None else p.nx * p.ny * p.nz

    def star_geometry(self):
        """(nb int64 [n, 4], w [n, 4], ds [n], shell_vol [n]) of the star set
        on this handle (host arrays)."""
        import numpy as np

        n = getattr(self, "_star_n", None)
        if n is None:
            raise TorchGpuError(1, "star_geometry: no star set")
        nb = np.zeros((n, 4), dtype=np.int64)
        w = np.zeros((n, 4))
        ds = np.zeros(n)
        sv = np.zeros(n)
        self._check(lib().torchgpu_star_geometry(self._h, nb.ctypes.data, w.ctypes.data, ds.ctypes.data,
                                                 sv.ctypes.data, 0))
        return nb, w, ds, sv

    def fill_heating(self, grid: "Grid", col_den=None, F_FUV=None, excluded: int = -1):
        """Thermodynamics::fillHeatingArrays on the state Torch::run prints:
        H [11][nz][ny][nx] (host, code units, HID order IMLC, NMLC, RHII,
        CEHI, CIEC, NMC, EUVH, FUVH, IRH, CRH, TOT).  The grid must track its
        heating terms (Grid.heating_track) from before the run."""
        import numpy as np

        p = grid.params
        H = np.zeros((11, p.nz, p.ny, p.nx))
        self._check(lib().torchgpu_fill_heating(grid._h, self._h, self._dptr(col_den, grid), self._dptr(F_FUV, grid),
                                                excluded, H.ctypes.data, 0))
        return H

    def grid_inputs(self, grid: "Grid", T_min=None, col_den=None, F_FUV=None, excluded: int = -1):
        """(in9 [9, n], DCOL_DEN [n]) the grid thermo operator forms for the
        current state (host arrays)."""
        import numpy as np

        n = grid.params.nx * grid.params.ny * grid.params.nz
        in9 = np.zeros((9, n))
        dcol = np.zeros(n)
        self._check(lib().torchgpu_thermo_grid_inputs(grid._h, self._h, self._dptr(T_min, grid), self._dptr(col_den, grid),
                                                      self._dptr(F_FUV, grid), excluded, in9.ctypes.data, dcol.ctypes.data,
                                                      0))
        return in9, dcol

    def thermo_cells(self, inp, out=None, count=True):
        """inp: device float64 tensor [9][n]; returns out [5][n] (device) and
        the total number of subcycles (if count)."""
        import torch

        if not inp.is_cuda or inp.dtype != torch.float64 or not inp.is_contiguous() or inp.dim() != 2 \
                or inp.shape[0] != 9:
            raise ValueError("inp must be a contiguous float64 CUDA tensor [9][n]")
        n = inp.shape[1]
        if out is None:
            out = torch.empty((5, n), dtype=torch.float64, device=inp.device)
        elif not out.is_cuda or out.dtype != torch.float64 or not out.is_contiguous() or tuple(out.shape) != (5, n):
            raise ValueError(f"out must be a contiguous float64 CUDA tensor [5][{n}]")
        tot = ctypes.c_int64()
        self._check(lib().torchgpu_thermo_cells(self._h, n, ctypes.c_void_p(inp.data_ptr()),
                                                ctypes.c_void_p(out.data_ptr()),
                                                ctypes.byref(tot) if count else None))
        return out, tot.value


# -- radiation (Radiation.cpp) -------------------------------------------------

class RadParams(ctypes.Structure):
    """torchgpu_rad_params (fill with rad_params)."""
    _fields_ = ([("K", ctypes.c_double * 4)]
                + [(k, ctypes.c_double) for k in ("sigma", "tau0", "min_hii", "T_hi", "T_hii")]
                + [(k, ctypes.c_int) for k in ("scheme", "coupling", "collisions_on", "pad")]
                + [(k, ctypes.c_double) for k in ("heating_amplification", "mass_fraction_h", "hydrogen_mass",
                                                  "spec_gas_const", "boltzmann", "rydberg", "pi", "voronov_A",
                                                  "voronov_X", "voronov_K", "voronov_P", "voronov_DE",
                                                  "photon_energy", "T_min")]
                + [("rec_x", ctypes.c_double * 31), ("rec_y", ctypes.c_double * 31), ("rec_y2", ctypes.c_double * 31),
                   ("rec_slope", ctypes.c_double * 2), ("cool_y", ctypes.c_double * 31),
                   ("cool_y2", ctypes.c_double * 31), ("cool_slope", ctypes.c_double * 2)])


RT_SCHEME = {"implicit": 0, "implicit2": 1, "explicit": 2}
RT_COUPLING = {"tti": 0, "neq": 1, "off": 2}


def _bind_radiation(L):
    if getattr(L, "_radiation_bound", False):
        return
    vp, d, i = ctypes.c_void_p, ctypes.c_double, ctypes.c_int
    L.torchgpu_rad_params_init.argtypes = [ctypes.POINTER(RadParams), d, d, d, ctypes.POINTER(d), d, d, d, d, d, i, i,
                                           i, d, d, d]
    L.torchgpu_radiation_create.argtypes = [ctypes.POINTER(RadParams), vp, ctypes.POINTER(Star), ctypes.POINTER(vp)]
    L.torchgpu_radiation_destroy.argtypes = [vp]
    L.torchgpu_radiation_destroy.restype = None
    L.torchgpu_radiation_init_field.argtypes = [vp, vp, i]
    L.torchgpu_radiation_timestep.argtypes = [vp, d, ctypes.POINTER(d)]
    L.torchgpu_radiation_pre_step.argtypes = [vp]
    L.torchgpu_radiation_substep.argtypes = [vp, d, i]
    L.torchgpu_radiation_get_state.argtypes = [vp, vp, vp, i]
    L.torchgpu_run_full.argtypes = [vp, vp, vp, vp, vp, vp, ctypes.c_int64, i, i]
    L._radiation_bound = True


def rad_params(scales=(1e-20, 1e-28, 1e11), K=(0.0, 0.0, 0.0, 0.0), photoion_cross_section=6.3e-18, tau0=0.6,
               min_hii_fraction=0.0, temperature_hi=300.0, temperature_hii=1e4, scheme="implicit", coupling="off",
               collisions_on=False, heating_amplification=1.0, mass_fraction_h=1.0,
               photon_energy=2.976e-11) -> RadParams:
    """Radiation::initialise from the Lua-level (cgs) values; unknown scheme
    and coupling strings fall back to implicit and off as the reference's."""
    L = lib()
    _bind_radiation(L)
    r = RadParams()
    Ka = (ctypes.c_double * 4)(*K)
    rc = L.torchgpu_rad_params_init(ctypes.byref(r), scales[0], scales[1], scales[2], Ka, photoion_cross_section,
                                    tau0, min_hii_fraction, temperature_hi, temperature_hii,
                                    RT_SCHEME.get(scheme, 0), RT_COUPLING.get(coupling, 2), int(collisions_on),
                                    heating_amplification, mass_fraction_h, photon_energy)
    if rc:
        raise TorchGpuError(rc, "rad_params_init")
    return r


class Radiation:
    """The Radiation integrator on a grid with a star (torchgpu_radiation)."""

    def __init__(self, params: RadParams, grid: "Grid", star: Star):
        L = lib()
        _bind_radiation(L)
        self.grid = grid
        self._h = ctypes.c_void_p()
        rc = L.torchgpu_radiation_create(ctypes.byref(params), grid._h, ctypes.byref(star), ctypes.byref(self._h))
        if rc:
            raise TorchGpuError(rc, ERRORS.get(rc, "radiation create failed"))

    def _check(self, rc):
        if rc:
            raise TorchGpuError(rc, ERRORS.get(rc, "radiation error"))

    def close(self):
        if self._h:
            lib().torchgpu_radiation_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def init_field(self, Q=None):
        """Radiation::initField; Q: the primitives given to set_primitive."""
        import numpy as np

        if Q is not None:
            Q = np.ascontiguousarray(Q, dtype=np.float64)
        self._check(lib().torchgpu_radiation_init_field(self._h, None if Q is None else ctypes.c_void_p(Q.ctypes.data),
                                                         0))

    def pre_step(self):
        """Q and Radiation::preTimeStepCalculations on the current state."""
        self._check(lib().torchgpu_radiation_pre_step(self._h))

    def timestep(self, dt_max: float) -> float:
        dt = ctypes.c_double()
        self._check(lib().torchgpu_radiation_timestep(self._h, dt_max, ctypes.byref(dt)))
        return dt.value

    def substep(self, dt: float, has_heat: bool):
        self._check(lib().torchgpu_radiation_substep(self._h, dt, int(has_heat)))

    def run_full(self, nsteps: int, cooling: "Cooling | None" = None, surrogate: bool = False, T_min=None,
                 col_den=None, F_FUV=None, excluded: int = -1):
        """`nsteps` Torch::fullStep iterations of hydro + (thermo) + radiation
        with dt, the pre-step calculations and the Strang alternation on the
        device (torchgpu_run_full, no host round trip)."""
        if cooling is None:
            dptr = lambda a: None  # noqa: E731
        else:
            dptr = lambda a: cooling._dptr(a, self.grid)  # noqa: E731
        self._check(lib().torchgpu_run_full(self.grid._h, None if cooling is None else cooling._h, self._h,
                                            dptr(T_min), dptr(col_den), dptr(F_FUV), excluded, nsteps,
                                            int(surrogate)))

    def state(self):
        """(R [6][nz][ny][nx], Q[HII, ADV] [2][nz][ny][nx]) host arrays."""
        import numpy as np

        p = self.grid.params
        shape = (p.nz, p.ny, p.nx)
        R = np.zeros((6,) + shape)
        q = np.zeros((2,) + shape)
        self._check(lib().torchgpu_radiation_get_state(self._h, ctypes.c_void_p(R.ctypes.data),
                                                        ctypes.c_void_p(q.ctypes.data), 0))
        return R, q
