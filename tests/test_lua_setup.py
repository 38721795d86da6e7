"""The Lua set-up interpreter (torch_amd/host/lua_setup.cpp) behind the
driver's --setupfile (Torch::setUpLua, Torch.cpp:192-228): set-up files are
run and their initialise(x, y, z, xs, ys, zs) is called for every cell with
the cell centre and the star position in cm.  Checked on the CPU through
--dump-setup against the same formulas evaluated by Python, bitwise (Python
floats and ** / math.* are the same C doubles and libm), on a file that
exercises the supported language, on a cloud set-up written like the
reference's config/torch-setup.lua, and -- where the reference is mounted --
on that file itself."""
import math
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "torch_amd", "torch_gpu")
GOLD = os.path.join(ROOT, "tests", "golden")
SCALES = (1.0e-20, 1.0e-28, 1.0e11)

PARAMS = """
Parameters = {
    Integration = { density_scale = 1.0e-20, pressure_scale = 1.0e-28, time_scale = 1.0e11,
                    spatial_order = 1, simulation_time = 1.0e12, ncheckpoints = 1 },
    Grid = { no_dimensions = 3, no_cells_x = 6, no_cells_y = 5, no_cells_z = 4, side_length = 3.09e18,
             geometry = "cartesian" },
    Hydrodynamics = { gamma = 1.67, density_floor = 1.0e-30, temperature_floor = 10 },
    Star = { on = true, cell_position_x = 2, cell_position_y = 3, cell_position_z = 1 },
}
"""


def _skip_without_driver():
    if not os.path.exists(EXE):
        pytest.skip("host driver not built (python -m torch_amd.build)")


def _dump(tmp_path, setup):
    pf = tmp_path / "p.lua"
    pf.write_text(PARAMS)
    out = tmp_path / "q.bin"
    r = subprocess.run([EXE, f"--paramfile={pf}", f"--setupfile={setup}", f"--dump-setup={out}"],
                       capture_output=True, text=True, timeout=60)
    return r, out


def _coords():
    """Cell centres and star position in cm as setUpLua passes them: index +
    0.5 times dx (code units), converted from code units."""
    import torch_amd as ta

    side = ta.units(SCALES, 3.09e18, 0, 1, 0)
    dx = side / 6
    back = lambda v: ta.units(SCALES, v, 0, 1, 0, to_code=False)  # noqa: E731
    xs = [back((s + 0.5) * dx) for s in (2, 3, 1)]
    cells = []
    for k in range(4):
        for j in range(5):
            for i in range(6):
                cells.append(([back((i + 0.5) * dx), back((j + 0.5) * dx), back((k + 0.5) * dx)], xs))
    return cells


def test_language_features(tmp_path):
    _skip_without_driver()
    r, out = _dump(tmp_path, os.path.join(GOLD, "setup_features.lua"))
    assert r.returncode == 0, r.stderr
    Q = np.fromfile(out).reshape(7, -1)
    checks = 2.0 ** 9 - 4 + 3 + 120 + (1 + 4 + 7 + 10 + 10 + 6 + 2) + 8 + 2 + 2 + 5 + 3 + 2 * 10 + 1 + 1000
    for c, (x, xs) in enumerate(_coords()):
        r2 = (x[0] - xs[0]) ** 2 + (x[1] - xs[1]) ** 2 + (x[2] - xs[2]) ** 2
        den = 1e-22 if r2 < 1e36 else (2e-22 if r2 < 4e36 else 3e-22)
        assert Q[0, c] == den
        assert Q[1, c] == den * 8.314462e7 * 300
        assert Q[2, c] == (1 if r2 < 1e35 else 0.25)
        assert Q[3, c] == 0.0  # ADV is not set by set-up files
        assert Q[4, c] == math.sqrt(r2) * 1e-18
        assert Q[5, c] == math.log10(r2 + 1) + 3 - 4
        assert Q[6, c] == checks + math.floor(-2.5) + abs(-1) + math.pow(2, 0.5) * math.cos(math.pi) + math.atan2(1, 1)


CLOUD = """
-- a cloud about 0.35 pc from the star along -y, in the manner of the
-- reference's example set-up
hydrogenMass = 1.674e-24
specificGasConstant = 8.314462e7
PC2CM = 3.09e18
T = 300.0
nHI = 32000.0
alpha = 1
rc = 0.01 * PC2CM
rc2 = rc * rc
RS = 0.35 * PC2CM
n0 = nHI * math.pow(1 + RS * RS / rc2, alpha)
pre0 = specificGasConstant * nHI * hydrogenMass * T

function initialise(x, y, z, xs, ys, zs)
    local dy = RS + (ys - y)
    local R2 = x * x + dy * dy
    local den = n0 * hydrogenMass * math.pow(1 + R2 / rc2, -alpha)
    return den, pre0, 0, 0, 0, 0, 0, 0, 0
end
"""


def _cloud_expect(x, xs):
    hm, R, pc = 1.674e-24, 8.314462e7, 3.09e18
    rc = 0.01 * pc
    rc2 = rc * rc
    RS = 0.35 * pc
    n0 = 32000.0 * math.pow(1 + RS * RS / rc2, 1)
    pre0 = R * 32000.0 * hm * 300.0
    dy = RS + (xs[1] - x[1])
    R2 = x[0] * x[0] + dy * dy
    return n0 * hm * math.pow(1 + R2 / rc2, -1), pre0


def test_cloud_setup_bitwise(tmp_path):
    _skip_without_driver()
    f = tmp_path / "cloud.lua"
    f.write_text(CLOUD)
    r, out = _dump(tmp_path, f)
    assert r.returncode == 0, r.stderr
    Q = np.fromfile(out).reshape(7, -1)
    for c, (x, xs) in enumerate(_coords()):
        den, pre = _cloud_expect(x, xs)
        assert Q[0, c] == den and Q[1, c] == pre


def test_reference_setup_file(tmp_path):
    """config/torch-setup.lua of the reference (read where mounted): the same
    cloud formulas."""
    _skip_without_driver()
    ref = "/root/reference/config/torch-setup.lua"
    if not os.path.exists(ref):
        pytest.skip("reference not mounted")
    r, out = _dump(tmp_path, ref)
    assert r.returncode == 0, r.stderr
    Q = np.fromfile(out).reshape(7, -1)
    for c, (x, xs) in enumerate(_coords()):
        den, pre = _cloud_expect(x, xs)
        assert Q[0, c] == den and Q[1, c] == pre
        assert (Q[2:, c] == 0).all()


@pytest.mark.parametrize("text,msg", [
    ("function initialise(x, y, z)\n  local a = (1 + \n  return a\nend\n", ":3:"),
    ("function initialise(x, y, z)\n  return math.random()\nend\n", "math.random"),
    ("x = 1\n", "initialise"),
    ("function initialise(x, y, z)\n  return 1, 2, 0, 0, 0, 0, 9.8, 0, 0\nend\n", "gravity"),
])
def test_errors_reported(tmp_path, text, msg):
    _skip_without_driver()
    f = tmp_path / "bad.lua"
    f.write_text(text)
    r, _ = _dump(tmp_path, f)
    assert r.returncode != 0 and msg in r.stderr, r.stderr
