"""The data2D snapshot format (DataPrinter::print2D, DataReader) against the
reference's own snapshot config/data2D_025.txt (tests/golden/data2D_025.txt.gz,
written by a 2-D run of the reference: 150 x 200 cells).  Host code only."""
import gzip
import os

import numpy as np
import pytest

import torch_amd as ta

FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "data2D_025.txt.gz")
SCALES = (1.0e-20, 1.0e-28, 1.0e11)


def lines(path):
    opener = gzip.open if path.endswith(".gz") else open
    with opener(path, "rt") as fh:
        return fh.read().split("\n")


def test_header():
    h = ta.snapshot_header(FIX)
    assert h["time"] == 1.577845e12 and h["shape"] == (1, 200, 150) and h["nd"] == 2
    assert h["side_length"] == pytest.approx(150 * 1.03e16, rel=1e-12)


def test_reference_snapshot_round_trip_is_byte_exact(tmp_path):
    """Read in cgs, written with unit scales 1: every line of the reference's
    file comes back byte for byte (the file lists cells rank by rank of an
    x-slab MPI run, the writer x fastest over the whole grid, so the lines
    are compared as a set; the header in order)."""
    Q, h = ta.snapshot_read(FIX)
    out = str(tmp_path / "out.txt")
    ta.snapshot_write(out, h["time"], Q, h["side_length"])
    a, b = lines(FIX), lines(out)
    assert a[:4] == b[:4]
    assert sorted(a[4:]) == sorted(b[4:])
    # x fastest: the second line is the cell right of the first
    assert b[5].split("\t")[0] == "1.5450000000e+16" and b[5].split("\t")[1] == "5.1500000000e+15"


def test_code_units_round_trip(tmp_path):
    """Through the reference's unit scalings (to code units and back, as a
    run reads and prints), gzip-compressed: every value within 2 units in
    the last printed digit."""
    Q, h = ta.snapshot_read(FIX)
    q = Q.copy()
    tc = lambda v, m, l, t: ta.units(SCALES, v, m, l, t)  # noqa: E731
    q[0] *= tc(1.0, 1, -3, 0)
    q[1] *= tc(1.0, 1, -1, -2)
    q[4:7] *= tc(1.0, 0, 1, -1)
    out = str(tmp_path / "out.txt.gz")
    ta.snapshot_write(out, tc(h["time"], 0, 0, 1), q, tc(h["side_length"], 0, 1, 0), scales=SCALES)
    Q2, h2 = ta.snapshot_read(out)
    assert h2["shape"] == h["shape"] and h2["time"] == pytest.approx(h["time"], rel=1e-10)
    for v in (0, 1, 2, 4, 5):
        np.testing.assert_allclose(Q2[v], Q[v], rtol=2e-10, atol=0)


def test_three_dimensional_state(tmp_path):
    """A 3-D state: cell centres (i + 1/2) dx, the velocity columns of all
    three axes, and reading places every cell by its coordinates."""
    rng = np.random.default_rng(5)
    Q = np.zeros((7, 3, 4, 5))
    Q[0] = rng.uniform(0.1, 2.0, Q.shape[1:])
    Q[1] = rng.uniform(0.1, 2.0, Q.shape[1:])
    Q[2] = rng.uniform(0.0, 1.0, Q.shape[1:])
    Q[4:7] = rng.normal(size=(3,) + Q.shape[1:])
    out = str(tmp_path / "s.txt")
    ta.snapshot_write(out, 0.25, Q, 1.0)
    ls = lines(out)
    assert ls[:4] == ["2.5000000000e-01", "5", "4", "3"]
    first = ls[4].split("\t")
    assert len(first) == 3 + 3 + 3 and first[:3] == ["1.0000000000e-01"] * 3
    Q2, h = ta.snapshot_read(out)
    assert h["nd"] == 3 and h["side_length"] == pytest.approx(1.0, rel=1e-12)
    for v in (0, 1, 2, 4, 5, 6):
        np.testing.assert_allclose(Q2[v], Q[v], rtol=1e-10, atol=1e-10)
    assert not Q2[3].any()  # ADV is not stored (lossy resume, as the reference)


def test_patch_averages_covered_cells(tmp_path):
    """DataReader::patchGrid: a 4 x 4 x 4 patch of half-width cells at offset
    (1, 2, 0) zeroes the 2 x 2 x 2 grid cells it covers and fills each with
    the mean of its 8 patch cells; the rest of the grid is untouched."""
    rng = np.random.default_rng(2)
    P = np.zeros((7, 4, 4, 4))
    P[0] = rng.uniform(1, 2, (4, 4, 4))
    P[1] = rng.uniform(1, 2, (4, 4, 4))
    P[2] = rng.uniform(0, 1, (4, 4, 4))
    P[4:7] = rng.normal(size=(3, 4, 4, 4))
    f = str(tmp_path / "patch.txt")
    ta.snapshot_write(f, 0.0, P, 4 * 0.5)  # patch dx 0.5
    P, _ = ta.snapshot_read(f)  # the values as printed (%.10e)
    Q = np.full((7, 3, 5, 6), 7.0)
    Q[3] = 0.25
    ta.snapshot_patch(f, Q, 1.0, offset=(1, 2, 0))
    blocks = P.reshape(7, 2, 2, 2, 2, 2, 2)
    for k in range(3):
        for j in range(5):
            for i in range(6):
                inside = 1 <= i < 3 and 2 <= j < 4 and 0 <= k < 2
                for v in (0, 1, 2, 4, 5, 6):
                    want = blocks[v, k, :, j - 2, :, i - 1, :].mean() if inside else 7.0
                    assert Q[v, k, j, i] == pytest.approx(want, rel=1e-15), (v, k, j, i)
                assert Q[3, k, j, i] == 0.25  # ADV is not patched


def test_patch_rejects_non_integer_ratio(tmp_path):
    P = np.ones((7, 4, 4, 4))
    f = str(tmp_path / "patch.txt")
    ta.snapshot_write(f, 0.0, P, 4 * 0.3)
    with pytest.raises(ta.TorchGpuError):
        ta.snapshot_patch(f, np.ones((7, 4, 4, 4)), 1.0)


def test_reference_file_patch():
    """The reference's 150 x 200 snapshot as a patch: at the file's own cell
    width it reproduces the file's cells (sum of one cell / 1); at twice the
    width it is refused, as the reference's check of the covered cell counts
    runs over all three axes and a 2-D file has one cell along z."""
    Qf, h = ta.snapshot_read(FIX)
    Q = np.full((7, 1, 200, 150), 3.0)
    ta.snapshot_patch(FIX, Q, 1.03e16)  # the file's dx, 1.03e16 cm
    for v in (0, 1, 2, 4, 5):
        np.testing.assert_array_equal(Q[v], Qf[v])
    assert (Q[3] == 3.0).all() and (Q[6] == 3.0).all()  # ADV and the third velocity untouched in 2-D
    with pytest.raises(ta.TorchGpuError):
        ta.snapshot_patch(FIX, np.zeros((7, 1, 100, 75)), 2 * 1.03e16)


def test_heating_text_matches_oracle_format(tmp_path):
    """torchgpu_heating_write against the oracle's restatement of
    DataPrinter::printHeating (DataPrinter.cpp:325-354): the header (time
    in s, nx, ny, nz), then per cell the raw cell-centre coordinates
    (GridCell::xc, index + 0.5) and the 11 terms in erg cm^-3 s^-1,
    "%.10e", tab separated, byte for byte.  The fixture's star run holds NaN
    terms where the radiation pre-step's rate is 0 (RHII and EUVH are
    x (soft / rate), Radiation.cpp:278-279); they print as the reference's
    stream prints them."""
    assert np.isnan(np.load(os.path.join(os.path.dirname(FIX), "heating.npz"))["star_H2"]).any()
    import oracle as orc

    f = np.load(os.path.join(os.path.dirname(FIX), "heating.npz"))
    H = f["star_H2"]
    t = float(f["star_time"])
    path = str(tmp_path / "heating_000001.txt.gz")
    ta.heating_write(path, t, H, SCALES)
    text = gzip.open(path, "rt").read()
    import math

    def fmt(v):  # printf's "%.10e", whose NaN keeps its sign
        if math.isnan(v):
            return "-nan" if math.copysign(1.0, v) < 0 else "nan"
        return f"{v:.10e}"

    c = orc.converter(*SCALES)
    nz, ny, nx = H.shape[1:]
    lines = [f"{orc.from_code(c, t, 0, 0, 1):.10e}", str(nx), str(ny), str(nz)]
    Hf = H.reshape(11, -1)
    for cidx in range(nx * ny * nz):
        ijk = (cidx % nx, (cidx // nx) % ny, cidx // (nx * ny))
        cols = [f"{a + 0.5:.10e}" for a in ijk]
        cols += [fmt(orc.from_code(c, Hf[i, cidx], 1, -1, -3)) for i in range(11)]
        lines.append("\t".join(cols))
    assert text == "\n".join(lines) + "\n"
