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
    assert d["debug"] is True  # TorchParameters default (Parameters.hpp:50)
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


def _sod16_lua(tmp_path, name, tmax, nck, outdir):
    cfg = open(os.path.join(ROOT, "torch_amd", "configs", "sod64.lua")).read()
    for k in ("x", "y", "z"):
        cfg = cfg.replace(f"no_cells_{k} = 64", f"no_cells_{k} = 16")
    cfg = cfg.replace("simulation_time = 1.0e3,", f"simulation_time = {tmax!r}, output_directory = \"{outdir}/\",")
    cfg = cfg.replace("ncheckpoints =    1,", f"ncheckpoints = {nck},")
    f = tmp_path / name
    f.write_text(cfg)
    return f


def _gunzip(path):
    import gzip

    with gzip.open(path, "rb") as fh:
        return fh.read()


@pytest.mark.gpu
def test_driver_snapshots_and_restart(tmp_path, orc, gpu):
    """--snapshots: the data2D files Torch::run prints (Torch.cpp:240-302) for
    16^3 Sod, tmax 0.1, 4 checkpoints, against the oracle's loop written by
    the same text writer (identical decompressed text).  Then a run started
    from the checkpoint-2 file as Integration.initial_conditions (grid shape,
    side length, state and time from the file, Torch.cpp:38-58,100-103)
    against the oracle started from the same values: same files, same final
    conserved state (bitwise)."""
    _exe_or_skip()
    from torch_amd import setups

    n, tmax, nck = 16, 0.1, 4
    f = _sod16_lua(tmp_path, "sod16.lua", tmax, nck, tmp_path / "out")
    r = subprocess.run([EXE, f"--paramfile={f}", "--setup=sod", "--steps=1000", "--snapshots"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    Q = setups.sod(n)
    po = orc.make_params(Q.shape[1:], 1.0 / n, dt_max=tmax / 10.0, nthreads=4)
    U0 = orc.UfromQ_grid(orc.fix_primitives(po, Q, np.zeros_like(Q)), po.gamma)
    Uo, dts, t, prints = orc.run_hydro_print(po, U0, 1000, tmax, nck)
    names = sorted(x for x in os.listdir(tmp_path / "out") if x.startswith("data2D"))
    assert names == [f"data2D_{c:06d}.txt.gz" for c, _, _ in prints] and len(names) == 5
    # printHeating before every print inside the loop (Torch.cpp:270-276), not
    # for the initial snapshot nor the final one after the loop (count
    # ncheckpoints, Torch.cpp:295-297): the oracle's fillHeatingArrays on the
    # oracle's printed state (the default HII switch is 0, so Sod's cells
    # count; cosmic-ray heating is their one nonzero term; the star-off
    # CausalWind cell (0, 0, 0) stays zero)
    heats = sorted(x for x in os.listdir(tmp_path / "out") if x.startswith("heating"))
    assert heats == [f"heating_{c:06d}.txt.gz" for c, _, _ in prints[1:] if c != nck] and len(heats) == 3
    tpar = orc.thermo_params(orc.converter(1.0, 1.0, 1.0), thermo_hii_switch=0.0)
    wind = np.zeros((n, n, n), dtype=bool)
    wind[0, 0, 0] = True
    for c, tp, Qp in prints:
        if c == 0 or c == nck:
            continue
        H = orc.fill_heating(tpar, Qp, np.zeros((3, n, n, n)), wind=wind)
        lines = _gunzip(tmp_path / "out" / f"heating_{c:06d}.txt.gz").decode().split("\n")
        assert lines[1:4] == ["16", "16", "16"] and len(lines) == 4 + n ** 3 + 1
        vals = np.array([[float(v) for v in line.split("\t")] for line in lines[4:-1]])
        np.testing.assert_array_equal(vals[:, :3], np.stack(np.meshgrid(np.arange(n) + 0.5, np.arange(n) + 0.5,
                                                                        np.arange(n) + 0.5, indexing="ij")[::-1],
                                                            -1).reshape(-1, 3))
        np.testing.assert_allclose(vals[:, 3:], H.reshape(11, -1).T, rtol=1e-9, atol=0)
        assert (vals[1:, 3 + 9] > 0).all() and (vals[0, 3:] == 0).all()
    for c, tp, Qp in prints:
        ref = str(tmp_path / f"oracle_{c}.txt.gz")
        gpu.snapshot_write(ref, tp, Qp, 1.0)
        assert _gunzip(tmp_path / "out" / f"data2D_{c:06d}.txt.gz") == _gunzip(ref), c

    # restart from checkpoint 2
    start = str(tmp_path / "out" / "data2D_000002.txt.gz")
    f2 = _sod16_lua(tmp_path, "restart.lua", tmax, nck, tmp_path / "out2")
    out = tmp_path / "U.bin"
    r = subprocess.run([EXE, f"--paramfile={f2}", f"--setup={start}", "--steps=1000", "--snapshots",
                        f"--output={out}"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    info = json.loads(r.stdout.strip().splitlines()[-1])
    Qs, h = gpu.snapshot_read(start)
    assert h["side_length"] == 1.0 and h["shape"] == (n, n, n)
    U1 = orc.UfromQ_grid(orc.fix_primitives(po, Qs, np.zeros_like(Qs)), po.gamma)
    Uo2, dts2, t2, prints2 = orc.run_hydro_print(po, U1, 1000, tmax, nck, t0=h["time"])
    assert [c for c, _, _ in prints2] == [2, 3, 4]
    assert sorted(x for x in os.listdir(tmp_path / "out2") if x.startswith("data2D")) == \
        [f"data2D_{c:06d}.txt.gz" for c, _, _ in prints2]
    for c, tp, Qp in prints2:
        ref = str(tmp_path / f"oracle2_{c}.txt.gz")
        gpu.snapshot_write(ref, tp, Qp, 1.0)
        assert _gunzip(tmp_path / "out2" / f"data2D_{c:06d}.txt.gz") == _gunzip(ref), c
    Ug = np.fromfile(out, dtype=np.float64).reshape(7, n, n, n)
    np.testing.assert_array_equal(Ug, Uo2)
    assert info["time"] == t2 and info["steps"] == int((dts2 > 0).sum())


COOL_LUA = """
Parameters = {
    Integration = {
        density_scale = 1.0e-20, pressure_scale = 1.0e-28, time_scale = 1.0e11,
        spatial_order = 1, temporal_order = 2, simulation_time = 1.0e13,
        radiation_on = false, cooling_on = true, ncheckpoints = 10,
    },
    Grid = {
        no_dimensions = 3, no_cells_x = 16, no_cells_y = 16, no_cells_z = 16,
        geometry = "cartesian", side_length = 1.54285e18,
        left_boundary_condition_x = "outflow", right_boundary_condition_x = "outflow",
        left_boundary_condition_y = "periodic", right_boundary_condition_y = "periodic",
        left_boundary_condition_z = "reflecting", right_boundary_condition_z = "reflecting",
    },
    Hydrodynamics = {
        gamma = 1.67, density_floor = 1.0e-30, temperature_floor = 0.1,
        riemann_solver = "RotatedHLLC", slope_limiter = "albada",
    },
    Radiation = { mass_fraction_hydrogen = 1.0 },
    Thermodynamics = { thermo_hii_switch = 1.0e-2, heating_amplification = 1.0 },
}
"""


@pytest.mark.gpu
def test_driver_cooling_matches_device_loop(tmp_path, gpu):
    """A 16^3 run with cooling_on through the driver's Torch::run with
    --host-loop (the host-driven Torch::fullStep sequence of the integrators)
    and, with --snapshots, through the device loop, against the library's device loop
    (torchgpu_run_split) driven from Python with the parameters the driver
    reports: identical conserved states.  Set-up: a warm (3000 K), partly
    ionised, turbulent medium of 1e-22 g/cm^3 with ADV = 1, so that both
    operators act."""
    _exe_or_skip()
    from torch_amd import setups

    n, steps = 16, 6
    f = tmp_path / "cool.lua"
    f.write_text(COOL_LUA)
    rng = np.random.default_rng(7)
    Qt = setups.turbulent(n, 0.5, 1.0)
    rho = 1.0e-22 * Qt[0]
    hii = rng.uniform(0.0, 1.0, rho.shape)
    Q = np.zeros((7, n, n, n))
    Q[0] = rho
    Q[1] = rho * 8.314462e7 * (1.0 + hii) * 3000.0
    Q[2] = hii
    Q[3] = 1.0
    Q[4:7] = Qt[4:7] / np.abs(Qt[4:7]).max() * 3.0e5
    setup = tmp_path / "Q.bin"
    Q.tofile(setup)
    d = json.loads(subprocess.run([EXE, f"--paramfile={f}", "--dump-params"], capture_output=True, text=True,
                                  timeout=60).stdout)
    res = {}
    for mode in ("host", "device"):
        out = tmp_path / f"U_{mode}.bin"
        cmd = [EXE, f"--paramfile={f}", f"--setup={setup}", f"--steps={steps}", f"--output={out}"]
        cmd.append("--snapshots" if mode == "device" else "--host-loop")
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=120, cwd=tmp_path)
        assert r.returncode == 0, r.stderr
        res[mode] = np.fromfile(out, dtype=np.float64).reshape(7, n, n, n)
        assert json.loads(r.stdout.strip().splitlines()[-1])["steps"] == steps
    U0out = tmp_path / "U0.bin"
    r = subprocess.run([EXE, f"--paramfile={f}", f"--setup={setup}", "--steps=0", f"--output={U0out}"],
                       capture_output=True, text=True, timeout=120, cwd=tmp_path)
    assert r.returncode == 0, r.stderr
    U0 = np.fromfile(U0out, dtype=np.float64).reshape(7, n, n, n)
    cu = d["code_units"]
    p = gpu.make_params((n, n, n), side_length=cu["side_length"], gamma=d["gamma"], dfloor=cu["dfloor"],
                        pfloor=cu["pfloor"], tfloor=d["tfloor"], spec_gas_const=cu["spec_gas_const"],
                        tmax=cu["tmax"], ncheckpoints=d["ncheckpoints"],
                        bc=["outflow", "outflow", "periodic", "periodic", "reflecting", "reflecting"])
    g = gpu.Grid(p)
    g.set_conserved(U0)
    tp = gpu.thermo_params(scales=tuple(d["scales"]), thermo_hii_switch=d["thermo_hii_switch"],
                           mass_fraction_h=d["mass_fraction_h"], gamma=d["gamma"], density_floor=d["dfloor"],
                           temperature_floor=d["tfloor"])
    tp.pfloor = cu["pfloor"]  # the driver hands the grid's pressure floor to the thermo operator
    cool = gpu.Cooling(tp)
    # the reference's thermo skips the CausalWind cell, here the star-off
    # default Star.cell_position (0, 0, 0)
    cool.run_split(g, steps, excluded=0)
    g.synchronize()
    Upy = g.get_conserved()
    assert not np.array_equal(Upy[1], U0[1])
    np.testing.assert_array_equal(res["device"], Upy)
    np.testing.assert_array_equal(res["host"], Upy)


STAR_LUA = """
Parameters = {
    Integration = {
        density_scale = 1.0e-20, pressure_scale = 1.0e-28, time_scale = 1.0e11,
        spatial_order = 1, temporal_order = 2, simulation_time = 1.0e13,
        radiation_on = true, cooling_on = true, ncheckpoints = 10,
    },
    Grid = {
        no_dimensions = 3, no_cells_x = 16, no_cells_y = 16, no_cells_z = 16,
        geometry = "cartesian", side_length = 1.0e18,
        left_boundary_condition_x = "outflow", right_boundary_condition_x = "outflow",
        left_boundary_condition_y = "outflow", right_boundary_condition_y = "outflow",
        left_boundary_condition_z = "outflow", right_boundary_condition_z = "outflow",
    },
    Hydrodynamics = {
        gamma = 1.67, density_floor = 1.0e-30, temperature_floor = 0.1,
        riemann_solver = "RotatedHLLC", slope_limiter = "albada",
    },
    Radiation = {
        K1 = 0.2, K2 = 0.0, K3 = 0.0, K4 = 0.0,
        photoion_cross_section = 6.3e-18, case_b_recombination_coeff = 2.59e-13, tau_0 = 0.6,
        minimum_hii_fraction = 0, temperature_hi = 300, temperature_hii = 10000,
        mass_fraction_hydrogen = 1.0, integration_scheme = "implicit", collisions_on = false, coupling = "neq",
    },
    Thermodynamics = { thermo_hii_switch = 1.0e-2, heating_amplification = 1.0 },
    Star = {
        on = true, cell_position_x = 7, cell_position_y = 8, cell_position_z = 6,
        snap_to_face_left_x = false, snap_to_face_left_y = false, snap_to_face_left_z = false,
        photon_energy = 2.976e-11, photon_rate = 4.9e48, wind_radius_in_cells = 1,
    },
}
"""


@pytest.mark.gpu
def test_driver_min_temp_initial_state(tmp_path, gpu):
    """Thermodynamics.min_temp_initial_state = true (the reference's shipped
    config): each cell's T_min is Fluid::calcTemperature of the set-up's cgs
    primitives with the code-unit gas constant, computed before
    Torch::toCodeUnits (Torch.cpp:113-117, Thermodynamics.cpp:102-111), so
    for these scales it sits ~1e-8 of the gas temperature.  The driver's run
    against the library's device loop given that field, and against the
    constant T_min = 100 K that `false` gives (cool 150-250 K gas: the soft
    landing below T_min + 200 K separates the two)."""
    _exe_or_skip()
    import torch

    n, steps = 16, 6
    rng = np.random.default_rng(11)
    rho = 1.0e-22 * 10.0 ** rng.uniform(-0.5, 0.5, (n, n, n))
    hii = rng.uniform(0.0, 0.05, rho.shape)
    T = rng.uniform(150.0, 250.0, rho.shape)
    Q = np.zeros((7, n, n, n))
    Q[0] = rho
    Q[1] = rho * 8.314462e7 * (1.0 + hii) * T
    Q[2] = hii
    Q[3] = 1.0
    setup = tmp_path / "Q.bin"
    Q.tofile(setup)
    res = {}
    for on in ("true", "false"):
        f = tmp_path / f"cool_{on}.lua"
        f.write_text(COOL_LUA.replace("heating_amplification = 1.0 }",
                                      f"heating_amplification = 1.0, min_temp_initial_state = {on} }}"))
        out = tmp_path / f"U_{on}.bin"
        r = subprocess.run([EXE, f"--paramfile={f}", f"--setup={setup}", f"--steps={steps}", f"--output={out}"],
                           capture_output=True, text=True, timeout=120, cwd=tmp_path)
        assert r.returncode == 0, r.stderr
        res[on] = np.fromfile(out, dtype=np.float64).reshape(7, n, n, n)
    d = json.loads(subprocess.run([EXE, f"--paramfile={f}", "--dump-params"], capture_output=True, text=True,
                                  timeout=60).stdout)
    U0out = tmp_path / "U0.bin"
    r = subprocess.run([EXE, f"--paramfile={f}", f"--setup={setup}", "--steps=0", f"--output={U0out}"],
                       capture_output=True, text=True, timeout=120, cwd=tmp_path)
    assert r.returncode == 0, r.stderr
    U0 = np.fromfile(U0out, dtype=np.float64).reshape(7, n, n, n)
    cu = d["code_units"]
    X = d["mass_fraction_h"]
    mu_inv = X * (Q[2] + 1.0) + (1.0 - X) * 0.25
    Tmin = (Q[1] / Q[0]) * (1.0 / mu_inv) / cu["spec_gas_const"]
    assert Tmin.max() < 1e-3 * T.min()
    ups = {}
    for key, field in (("true", Tmin), ("false", None)):
        p = gpu.make_params((n, n, n), side_length=cu["side_length"], gamma=d["gamma"], dfloor=cu["dfloor"],
                            pfloor=cu["pfloor"], tfloor=d["tfloor"], spec_gas_const=cu["spec_gas_const"],
                            tmax=cu["tmax"], ncheckpoints=d["ncheckpoints"],
                            bc=["outflow", "outflow", "periodic", "periodic", "reflecting", "reflecting"])
        g = gpu.Grid(p)
        g.set_conserved(U0)
        tp = gpu.thermo_params(scales=tuple(d["scales"]), thermo_hii_switch=d["thermo_hii_switch"],
                               mass_fraction_h=X, gamma=d["gamma"], density_floor=d["dfloor"],
                               temperature_floor=d["tfloor"])
        tp.pfloor = cu["pfloor"]
        cool = gpu.Cooling(tp)
        tm = None if field is None else torch.from_numpy(np.ascontiguousarray(field).ravel()).cuda()
        cool.run_split(g, steps, T_min=tm, excluded=0)
        g.synchronize()
        ups[key] = g.get_conserved()
    np.testing.assert_array_equal(res["true"], ups["true"])
    np.testing.assert_array_equal(res["false"], ups["false"])
    assert not np.array_equal(res["true"][1], res["false"][1])


@pytest.mark.gpu
def test_driver_star_radiation_cooling_matches_python_loop(tmp_path, gpu):
    """Torch::fullStep with hydrodynamics, cooling and radiation around a star
    (Lua Star and Radiation tables): the C++ driver's host-driven sequence
    against torch_amd.fullstep.FullStep over the same library calls with the
    parameters the driver reports: identical conserved states after 5 steps
    (the first with dt_max 1e-20, then min(hydro, radiation, thermo))."""
    _exe_or_skip()
    from torch_amd import fullstep, setups

    n, steps = 16, 5
    f = tmp_path / "star.lua"
    f.write_text(STAR_LUA)
    d = json.loads(subprocess.run([EXE, f"--paramfile={f}", "--dump-params"], capture_output=True, text=True,
                                  timeout=60).stdout)
    assert d["star_on"] and d["radiation_on"] and d["star_cell"] == [7, 8, 6] and d["coupling"] == "neq"
    Qt = setups.turbulent(n, 0.5, 0.5)
    rho = 1.0e-21 * Qt[0]
    Q = np.zeros((7, n, n, n))
    Q[0] = rho
    Q[1] = rho * 8.314462e7 * 100.0
    Q[3] = 1.0
    Q[4:7] = Qt[4:7] / np.abs(Qt[4:7]).max() * 1.0e5
    setup = tmp_path / "Q.bin"
    Q.tofile(setup)
    outs = {}
    for k in (0, steps):
        out = tmp_path / f"U_{k}.bin"
        r = subprocess.run([EXE, f"--paramfile={f}", f"--setup={setup}", f"--steps={k}", f"--output={out}"],
                           capture_output=True, text=True, timeout=120, cwd=tmp_path)
        assert r.returncode == 0, r.stderr
        outs[k] = np.fromfile(out, dtype=np.float64).reshape(7, n, n, n)
    scales = tuple(d["scales"])
    cu = d["code_units"]
    p = gpu.make_params((n, n, n), side_length=cu["side_length"], gamma=d["gamma"], dfloor=cu["dfloor"],
                        pfloor=cu["pfloor"], tfloor=d["tfloor"], spec_gas_const=cu["spec_gas_const"],
                        tmax=cu["tmax"], ncheckpoints=d["ncheckpoints"])
    g = gpu.Grid(p)
    g.set_conserved(outs[0])
    star = gpu.make_star(d["star_cell"], d["wind_radius"], cu["photon_rate"])
    tp = gpu.thermo_params(scales=scales, thermo_hii_switch=d["thermo_hii_switch"],
                           mass_fraction_h=d["mass_fraction_h"], gamma=d["gamma"], density_floor=d["dfloor"],
                           temperature_floor=d["tfloor"])
    tp.pfloor = cu["pfloor"]
    cool = gpu.Cooling(tp)
    cool.set_star(g, star)
    rp = gpu.rad_params(scales, K=tuple(d["K"]), photoion_cross_section=d["photoion_cross_section"],
                        tau0=d["tau_0"], temperature_hi=d["temperature_hi"], temperature_hii=d["temperature_hii"],
                        scheme=d["integration_scheme"], coupling=d["coupling"], collisions_on=d["collisions_on"],
                        heating_amplification=d["heating_amplification"], mass_fraction_h=d["mass_fraction_h"],
                        photon_energy=d["photon_energy"])
    rad = gpu.Radiation(rp, g, star)
    # the code-unit primitives Torch::initialise hands to Radiation::initField
    to = np.vectorize(lambda v, m, l, t: gpu.units(scales, v, m, l, t))
    Qc = Q.copy()
    Qc[0] = to(Q[0], 1, -3, 0)
    Qc[1] = to(Q[1], 1, -1, -2)
    for a in range(3):
        Qc[4 + a] = to(Q[4 + a], 0, 1, -1)
    rad.init_field(Qc)
    loop = fullstep.FullStep(g, cooling=cool, radiation=rad)
    dts = [loop.step() for _ in range(steps)]
    g.synchronize()
    U = g.get_conserved()
    assert np.isfinite(U).all() and dts[1] > dts[0]
    assert not np.array_equal(U[2], outs[0][2])  # the star ionised gas
    np.testing.assert_array_equal(outs[steps], U)
    # the driver's host-driven Torch::fullStep (--host-loop) and its device
    # loop with checkpoint printing (--snapshots) give the same state
    for flag in ("--host-loop", "--snapshots"):
        out = tmp_path / f"U_{flag[2:]}.bin"
        r = subprocess.run([EXE, f"--paramfile={f}", f"--setup={setup}", f"--steps={steps}", f"--output={out}", flag],
                           capture_output=True, text=True, timeout=120, cwd=tmp_path)
        assert r.returncode == 0, (flag, r.stderr)
        np.testing.assert_array_equal(np.fromfile(out, dtype=np.float64).reshape(7, n, n, n), U, err_msg=flag)


@pytest.mark.gpu
def test_driver_lua_setupfile_run(tmp_path):
    """--setupfile=FILE.lua (Torch::setUpLua) drives a run: 4 steps from the
    cloud set-up equal 4 steps from the primitives the same file produces
    through --dump-setup (bitwise)."""
    _exe_or_skip()
    from test_lua_setup import CLOUD, PARAMS

    pf = tmp_path / "p.lua"
    pf.write_text(PARAMS)
    sf = tmp_path / "cloud.lua"
    sf.write_text(CLOUD)
    qbin = tmp_path / "q.bin"
    r = subprocess.run([EXE, f"--paramfile={pf}", f"--setupfile={sf}", f"--dump-setup={qbin}"], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    outs = []
    for setup in (f"--setupfile={sf}", f"--setup={qbin}"):
        out = tmp_path / f"U{len(outs)}.bin"
        r = subprocess.run([EXE, f"--paramfile={pf}", setup, "--steps=4", f"--output={out}"], capture_output=True,
                           text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        outs.append(np.fromfile(out))
    assert np.isfinite(outs[0]).all()
    np.testing.assert_array_equal(outs[0], outs[1])


@pytest.mark.gpu
@pytest.mark.parametrize("host_loop", [False, True])
@pytest.mark.parametrize("debug", [None, False])
def test_driver_debug_stop(tmp_path, orc, host_loop, debug):
    """Integration.debug through the driver: a 16^3 Sod run with
    simulation_time 1e6 has 100 dt_hydro / tmax below 1e-6 at its first CFL
    step, so Torch::run ends after step 2 (Torch.cpp:321-332) -- both the
    device loop and the host-driven fullStep; with debug = false it runs all
    its steps.  The state equals the oracle's run with the same rule."""
    _exe_or_skip()
    from torch_amd import setups

    cfg = open(os.path.join(ROOT, "torch_amd", "configs", "sod64.lua")).read()
    for k in ("x", "y", "z"):
        cfg = cfg.replace(f"no_cells_{k} = 64", f"no_cells_{k} = 16")
    cfg = cfg.replace("simulation_time = 1.0e3", "simulation_time = 1.0e6")
    if debug is not None:
        cfg = cfg.replace("ncheckpoints =    1,", f"ncheckpoints =    1, debug = {str(debug).lower()},")
    f = tmp_path / "sod16.lua"
    f.write_text(cfg)
    out = tmp_path / "U.bin"
    cmd = [EXE, f"--paramfile={f}", "--setup=sod", "--steps=6", f"--output={out}"]
    if host_loop:
        cmd.append("--host-loop")
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    info = json.loads(r.stdout.strip().splitlines()[-1])
    stop = debug is None
    assert info["steps"] == (2 if stop else 6) and info["quitting"] is stop
    assert ("Integration deltas are too small." in r.stderr) is stop
    Q = setups.sod(16)
    p = orc.make_params((16, 16, 16), 1.0 / 16, dt_max=1e5, nthreads=8)
    U0 = orc.UfromQ_grid(orc.fix_primitives(p, Q, np.zeros_like(Q)), p.gamma)
    Uo, dts, t, quit_ = orc.run_hydro(p, U0, 6, tmax=1e6, debug=stop, info=True)
    assert quit_ is stop and len(dts) == info["steps"]
    np.testing.assert_array_equal(np.fromfile(out, dtype=np.float64).reshape(7, 16, 16, 16), Uo)
