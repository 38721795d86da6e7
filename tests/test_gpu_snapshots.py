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


def test_mixed_run_prints_the_predictor_state_it_skips_storing(gpu, tmp_path):
    """The mixed predictor stores its FP64 U* only where something reads it:
    the last step of a run call, and a step after which the device loop can
    print (it reaches a new checkpoint or tmax).  A run that pauses at
    checkpoints in one call must print exactly what a run of one-step calls
    prints (every one of those steps stores U*): the same files, the same
    final state.  Turbulent 24^3, 4 checkpoints, mixed path."""
    n, tmax, nck = 24, 0.05, 4
    Q0 = setups.turbulent(n, 1.0, 2.0)
    kw = dict(side_length=1.0, tmax=tmax, ncheckpoints=nck, tfloor=1e-12, pfloor=0.1 * 8.314462e7 * 1e-30 * 1e-12,
              precision=gpu.MIXED)
    ga = gpu.Grid(gpu.make_params((n, n, n), **kw))
    ga.set_primitive(Q0)
    files_a = gpu.run_printing(ga, 100000, str(tmp_path / "a"))
    ga.synchronize()

    gb = gpu.Grid(gpu.make_params((n, n, n), **kw))
    gb.set_primitive(Q0)
    prints_b = [(0, gb.clock()[0], gb.print_state())]
    gb.pause_at_checkpoints(True)
    final = False
    for _ in range(100000):
        gb.run(1)
        gb.synchronize()
        count, due = gb.checkpoint()
        if due:
            prints_b.append((count, gb.clock()[0], gb.print_state()))
            final = count != nck
        elif gb.clock()[0] >= tmax:
            break
    if final:
        prints_b.append((nck, gb.clock()[0], gb.print_state()))
    gb.pause_at_checkpoints(False)
    assert len(files_a) == len(prints_b) == nck + 1, (files_a, [c for c, _, _ in prints_b])
    for f, (c, t, Qb) in zip(files_a, prints_b):
        Qa, h = gpu.snapshot_read(f)
        ref = str(tmp_path / f"b_{c}.txt.gz")
        gpu.snapshot_write(ref, t, Qb, 1.0)
        Qr, hr = gpu.snapshot_read(ref)
        assert h["time"] == hr["time"], c
        np.testing.assert_array_equal(Qa, Qr, err_msg=f"checkpoint {c}")
    np.testing.assert_array_equal(ga.get_conserved(), gb.get_conserved())
    assert ga.clock() == gb.clock()
