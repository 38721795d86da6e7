"""The C++ host layer (torch_amd/host): the Lua parameter reader and the
Torch::run mirror driving the library.

CPU: `torch_gpu --dump-params` parses a parameter file written with the
reference's keys (Parameters.hpp defaults, main.cpp:135-204) and reports the
values TorchParameters holds after initialise (code units).
GPU: a 16^3 Sod run through the driver equals the oracle bitwise (same
steps, same state)."""
import json
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "torch_amd", "torch_gpu")

PARAMS = """
-- reference key names; arithmetic and a local alias as in the reference's configs
local cells = 16
Parameters = {
    Integration = {
        density_scale = 1.0e-20, pressure_scale = 1.0e-28, time_scale = 1.0e11,
        spatial_order = 1, temporal_order = 2, simulation_time = 2.0e12,
        radiation_on = false, cooling_on = false, ncheckpoints = 4,
    },
    Grid = {
        no_dimensions = 3, no_cells_x = cells, no_cells_y = cells / 2, no_cells_z = 2 * 4,
        geometry = "cartesian", side_length = 3.0857e18,
        left_boundary_condition_x = "reflecting", right_boundary_condition_x = 'outflow',
        left_boundary_condition_y = "periodic", right_boundary_condition_y = "periodic",
        left_boundary_condition_z = "free", right_boundary_condition_z = "inflow",
    },
    Hydrodynamics = {
        gamma = 5.0 / 3.0, density_floor = 1.0e-30, temperature_floor = 10,
        riemann_solver = "HLLC", slope_limiter = "minmod", precision = "mixed",
    },
    Radiation = { mass_fraction_hydrogen = 0.7 },
    Thermodynamics = { thermo_hii_switch = 0.5 },
}
"""


def _exe_or_skip():
    if not os.path.exists(EXE):
        pytest.skip("host driver not built (python -m torch_amd.build)")


def test_lua_parameters_parsed(tmp_path):
    _exe_or_skip()
    f = tmp_path / "p.lua"
    f.write_text(PARAMS)
    r = subprocess.run([EXE, f"--paramfile={f}", "--dump-params"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    d = json.loads(r.stdout)
    assert d["ncells"] == [16, 8, 8] and d["nd"] == 3 and d["spatial_order"] == 1
    assert d["riemann_solver"] == "HLLC" and d["slope_limiter"] == "minmod" and d["precision"] == "mixed"
    assert d["bc"] == ["reflecting", "outflow", "periodic", "periodic", "free", "inflow"]
    assert d["gamma"] == 5.0 / 3.0 and d["mass_fraction_h"] == 0.7 and d["thermo_hii_switch"] == 0.5
    assert d["scales"] == [1e-20, 1e-28, 1e11] and d["ncheckpoints"] == 4 and d["cooling_on"] is False
    assert d["side_length"] == 3.0857e18 and d["tmax"] == 2.0e12 and d["tfloor"] == 10.0
    # after TorchParameters::initialise: L = sqrt(p/rho) t = 1e7 cm, T = 1e11 s,
    # M = L^3 rho = 1e1 g (Converter.cpp:22-24); dt_max = tmax / 10, pfloor =
    # 0.1 R dfloor tfloor (Parameters.cpp:9-27)
    cu = d["code_units"]
    assert abs(cu["side_length"] - 3.0857e11) < 1e-9 * 3.0857e11
    assert abs(cu["tmax"] - 20.0) < 1e-12 and abs(cu["dt_max"] - 2.0) < 1e-12
    assert abs(cu["dfloor"] - 1e-10) < 1e-22
    R_code = 8.314462e7 / (1e7 / 1e11) ** 2  # erg/(g K) = (cm/s)^2 / K -> (L/T)^2 / K
    assert abs(cu["pfloor"] - 0.1 * R_code * 1e-10 * 10.0) < 1e-9 * cu["pfloor"]


def test_lua_errors_reported(tmp_path):
    _exe_or_skip()
    f = tmp_path / "bad.lua"
    f.write_text("Parameters = { Grid = { no_cells_x = 16 ")
    r = subprocess.run([EXE, f"--paramfile={f}", "--dump-params"], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "line" in r.stderr


@pytest.mark.gpu
def test_driver_sod_bitwise(tmp_path, orc):
    """torch_amd/configs/sod64.lua with 16^3 cells through Torch::run on the
    GPU against the oracle's run_hydro: identical conserved state."""
    _exe_or_skip()
    from torch_amd import setups

    cfg = open(os.path.join(ROOT, "torch_amd", "configs", "sod64.lua")).read()
    for k in ("x", "y", "z"):
        cfg = cfg.replace(f"no_cells_{k} = 64", f"no_cells_{k} = 16")
    f = tmp_path / "sod16.lua"
    f.write_text(cfg)
    out = tmp_path / "U.bin"
    r = subprocess.run([EXE, f"--paramfile={f}", "--setup=sod", "--steps=6", f"--output={out}"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    info = json.loads(r.stdout.strip().splitlines()[-1])
    assert info["steps"] == 6
    Ug = np.fromfile(out, dtype=np.float64).reshape(7, 16, 16, 16)
    Q = setups.sod(16)
    p = orc.make_params(Q.shape[1:], 1.0 / 16, nthreads=4)
    U0 = orc.UfromQ_grid(orc.fix_primitives(p, Q, np.zeros_like(Q)), p.gamma)
    Uo, dts, t = orc.run_hydro(p, U0, 6)
    np.testing.assert_array_equal(Ug, Uo)
    assert info["time"] == t
