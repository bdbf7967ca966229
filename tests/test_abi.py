"""The C-ABI library loads and exports every entry point include/mhd_api.h
declares; host-only paths (argument checks, file I/O) work without a GPU."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import GOLDEN, ROOT, bitwise

HEADER = os.path.join(ROOT, "include", "mhd_api.h")


def declared():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(mhd_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_entry_points():
    names = declared()
    for must in ("mhd_fluidx", "mhd_advectbyzx", "mhd_transpose12", "mhd_transpose13", "mhd_cfl", "mhd_init",
                 "mhd_main", "mhd_ctx_create", "mhd_ctx_step"):
        assert must in names


def test_library_exports_every_declared_symbol():
    import mhdgpu
    L = mhdgpu.lib()
    missing = [n for n in declared() if not hasattr(L, n)]
    assert not missing, missing
    # and the Python mirror binds every one of them
    assert not [n for n in declared() if n not in mhdgpu.SIGNATURES]
    assert mhdgpu.lib().mhd_version().decode().startswith("mhdgpu")


def test_argument_errors_are_codes_not_exits():
    import mhdgpu
    L = mhdgpu.lib()
    n = ctypes.c_int(16)
    assert L.mhd_fluidx(None, None, ctypes.byref(n), ctypes.byref(n), ctypes.byref(n), None) == -1
    assert "null" in L.mhd_last_error().decode()
    p = ctypes.c_void_p()
    # pencils above 1024 cells (the reference's one-block limit), and FP64
    # pencils above 512, are refused before any device work
    assert L.mhd_ctx_create(ctypes.byref(p), 1025, 8, 8, 4, 0) == -2
    assert L.mhd_ctx_create(ctypes.byref(p), 8, 600, 8, 8, 0) == -2
    assert L.mhd_ctx_create(ctypes.byref(p), 8, 8, 8, 3, 0) == -1
    # the batched form takes only the fast kernels' lengths (n = L*K, K <= 8)
    dims = (ctypes.c_int * 3)(67, 8, 8)
    assert L.mhd_batch_create(ctypes.byref(p), 1, dims, 0) == -2
    with pytest.raises(mhdgpu.MHDError) as e:
        mhdgpu.init(os.path.join(GOLDEN, "does_not_exist.dat"), 16)
    assert e.value.code == -6


def test_init_reader_and_result_writer(oracle, tmp_path):
    import mhdgpu
    path = os.path.join(GOLDEN, "source_init_alfvenlinear_16cube.dat")
    u, b = mhdgpu.init(path, 16, order=0)
    rec = np.fromfile(path, dtype="<f4")
    uo, bo = oracle.read_ref_order(rec, 16)
    assert bitwise(u, uo) and bitwise(b, bo)
    # the result writer uses the reader's record order (mhd.c:130-161)
    pu, pb = tmp_path / "ru.dat", tmp_path / "rb.dat"
    mhdgpu.write_result(pu, pb, u, b, 16, 16, 16)
    ru = np.fromfile(pu, dtype="<f4").reshape(-1, 5)
    rb = np.fromfile(pb, dtype="<f4").reshape(-1, 3)
    assert bitwise(np.concatenate([ru, rb], 1).ravel(), rec)
    # plain Fortran order
    u1, b1 = mhdgpu.init(path, 16, order=1)
    assert bitwise(u1.reshape(-1, 5), rec.reshape(-1, 8)[:, :5])
