"""Torch::run's snapshots (Torch.cpp:240-302: the initial state, one per
checkpoint, the final one) from the device run with its checkpoint pauses,
against the oracle's restatement of the same loop: the same checkpoint
counts and times, and the same printed primitive state -- which after a
step is the reference's stale Fluid::Q, QfromU(U*) of the last corrector.
Sod 16^3, tmax 0.1, 4 checkpoints, FP64 path (bitwise)."""
import os

import numpy as np
import pytest

from torch_amd import setups

pytestmark = pytest.mark.gpu


def test_checkpoint_snapshots_match_oracle(gpu, orc, tmp_path):
    n, tmax, nck = 16, 0.1, 4
    Q0 = setups.sod(n)
    g = gpu.Grid(gpu.make_params((n, n, n), side_length=1.0, tmax=tmax, ncheckpoints=nck))
    g.set_primitive(Q0)
    files = gpu.run_printing(g, 1000, str(tmp_path / "gpu"))
    g.synchronize()
    po = orc.make_params((n, n, n), 1.0 / n, dt_max=tmax / 10.0, nthreads=8)
    U0 = orc.UfromQ_grid(orc.fix_primitives(po, Q0, np.zeros_like(Q0)), po.gamma)
    Uo, dts, t, prints = orc.run_hydro_print(po, U0, 1000, tmax, nck)
    assert [os.path.basename(f) for f in files] == [f"data2D_{c:06d}.txt.gz" for c, _, _ in prints]
    assert [c for c, _, _ in prints] == [0, 1, 2, 3, 4]
    tg, _, steps = g.clock()
    assert tg == t and steps == int((dts > 0).sum())
    np.testing.assert_array_equal(g.get_conserved(), Uo)
    # the final printed state bitwise, every file's text through the same writer
    np.testing.assert_array_equal(g.print_state(), prints[-1][2])
    for f, (c, tp, Qp) in zip(files, prints):
        ref = str(tmp_path / f"oracle_{c}.txt.gz")
        gpu.snapshot_write(ref, tp, Qp, 1.0)
        Qg, hg = gpu.snapshot_read(f)
        Qr, hr = gpu.snapshot_read(ref)
        assert hg["time"] == hr["time"], c
        np.testing.assert_array_equal(Qg, Qr, err_msg=f"checkpoint {c}")
