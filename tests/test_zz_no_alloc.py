"""torchgpu_thermo_cells_noalloc (the subcycled cooling step on independent
cells with its work-order scratch in the library's fixed device array, no
allocation) against torchgpu_thermo_cells, through the C ABI: bitwise equal
outputs and subcycle totals on two calls in a row, the second with fewer
cells, and the size limit refused before any launch."""
import ctypes
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
MAX_CELLS = 1 << 20


def _bind(L):
    vp = ctypes.c_void_p
    L.torchgpu_thermo_cells_noalloc.argtypes = [vp, ctypes.c_int64, vp, vp, ctypes.POINTER(ctypes.c_int64)]
    L.torchgpu_thermo_cells.argtypes = [vp, ctypes.c_int64, vp, vp, ctypes.POINTER(ctypes.c_int64)]
    return L


def test_noalloc_equals_thermo_cells(gpu):
    import torch

    L = _bind(gpu.lib())
    f = np.load(os.path.join(GOLD, "thermo.npz"))
    inp_all = np.ascontiguousarray(f["inputs"].T)   # [9][n]
    tp = gpu.thermo_params()
    tp.pfloor = float(f["pfloor"])
    cool = gpu.Cooling(tp)
    for n in (inp_all.shape[1], 1500):
        inp = torch.from_numpy(np.ascontiguousarray(inp_all[:, :n])).cuda()
        out_ref = torch.zeros((5, n), dtype=torch.float64, device="cuda")
        out_na = torch.full((5, n), float("nan"), dtype=torch.float64, device="cuda")
        tot_ref, tot_na = ctypes.c_int64(-1), ctypes.c_int64(-1)
        torch.cuda.synchronize()
        rc = L.torchgpu_thermo_cells(cool._h, n, inp.data_ptr(), out_ref.data_ptr(), ctypes.byref(tot_ref))
        assert rc == 0, rc
        rc = L.torchgpu_thermo_cells_noalloc(cool._h, n, inp.data_ptr(), out_na.data_ptr(), ctypes.byref(tot_na))
        assert rc == 0, rc
        torch.cuda.synchronize()
        assert tot_na.value == tot_ref.value and tot_ref.value > n
        np.testing.assert_array_equal(out_na.cpu().numpy(), out_ref.cpu().numpy())
    cool.close()


def test_noalloc_size_limit(gpu):
    import torch

    L = _bind(gpu.lib())
    cool = gpu.Cooling(gpu.thermo_params())
    buf = torch.zeros((9, 4), dtype=torch.float64, device="cuda")
    rc = L.torchgpu_thermo_cells_noalloc(cool._h, MAX_CELLS + 1, buf.data_ptr(), buf.data_ptr(), None)
    assert rc == 1
    msg = ctypes.create_string_buffer(256)
    gpu.lib().torchgpu_cooling_last_error(cool._h, msg, 256)
    assert b"at most 1048576" in msg.value
    cool.close()
