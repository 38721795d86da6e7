"""The C-ABI library (torch_amd/libtorchgpu.so) loads on a machine without a
GPU and exports every function include/torchgpu.h declares; its argument
checks reject invalid handles before any device work."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "torchgpu.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    # declarations start at column 0 with a return type
    return sorted(set(re.findall(r"^(?:int|void)\s+\**(torchgpu_\w+)\s*\(", text, flags=re.M)))


@pytest.fixture(scope="module")
def lib():
    from torch_amd import LIB_PATH

    if not os.path.exists(LIB_PATH):
        from torch_amd import build

        build.build()
    return ctypes.CDLL(LIB_PATH)


def test_header_declares_the_interface():
    names = declared_functions()
    assert len(names) >= 30, names
    for n in ("torchgpu_create", "torchgpu_hydro_step", "torchgpu_run", "torchgpu_thermo_cells",
              "torchgpu_thermo_cells_surrogate", "torchgpu_mlp_train_step", "torchgpu_thermo_grid_substep"):
        assert n in names


def test_every_declared_symbol_is_exported(lib):
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_null_handles_rejected_without_device_work(lib):
    from torch_amd import ERRORS

    invalid = [k for k, v in ERRORS.items() if "invalid" in v.lower()]
    assert invalid
    code = invalid[0]
    vp = ctypes.c_void_p
    lib.torchgpu_hydro_step.argtypes = [vp, ctypes.c_double]
    lib.torchgpu_synchronize.argtypes = [vp]
    lib.torchgpu_thermo_cells.argtypes = [vp, ctypes.c_int64, vp, vp, vp]
    lib.torchgpu_create.argtypes = [vp, vp, vp]
    assert lib.torchgpu_hydro_step(None, 1.0) == code
    assert lib.torchgpu_synchronize(None) == code
    assert lib.torchgpu_thermo_cells(None, 1, None, None, None) == code
    assert lib.torchgpu_create(None, None, None) == code
