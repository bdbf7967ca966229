"""The boundary as a C and a Fortran program link it (tests/drivers/, built
against include/mhd_api.h by __graft_entry__.build()):

* drop_in_main.c is the reference's program (HIP/mhd.cpp:39-64: init, then
  cuda_main on host arrays) on this library: the 16^3 fixture, 10 steps; its
  result files (Heterogeneous_MHD/C/mhd.c:126-163 record order) and dt
  sequence must equal the golden vectors the reference build wrote;
* transpose_f.f90 calls the Fortran-callable host-array transposes with the
  reference's calling convention (cuda_transpose12_, HIP/cu/cuda_transpose12.cu:
  69-127) and checks the component maps and the involution itself;
* the same host-array transposes through ctypes against the oracle.
"""
import os
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT, bitwise, same
from golden.states import random_state

pytestmark = pytest.mark.gpu
BIN = os.path.join(ROOT, "tests", "drivers", "bin")


def driver(name):
    exe = os.path.join(BIN, name)
    assert os.path.exists(exe), f"{exe} missing: run __graft_entry__.build() (make -C tests/drivers)"
    return exe


def from_records(rec, ncomp, n):
    """Result-file records (i outermost, k innermost) to the AoS array."""
    return np.ascontiguousarray(rec.reshape(n, n, n, ncomp).transpose(2, 1, 0, 3)).ravel()


def test_c_program_drop_in_equals_reference_golden(golden, tmp_path):
    fixture = os.path.join(GOLDEN, "source_init_alfvenlinear_16cube.dat")
    r = subprocess.run([driver("drop_in_main"), fixture, "16", "10", str(tmp_path)], capture_output=True, text=True,
                       timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stderr
    u = from_records(np.fromfile(tmp_path / "result_u1.dat", dtype="<f4"), 5, 16)
    b = from_records(np.fromfile(tmp_path / "result_b1.dat", dtype="<f4"), 3, 16)
    dts = np.fromfile(tmp_path / "dts.bin", dtype="<f4")
    assert bitwise(dts, golden["fix16_run10_dt"])
    assert same(u, golden["fix16_run10_u"]) and same(b, golden["fix16_run10_b"])


def test_fortran_program_calls_the_host_transposes():
    r = subprocess.run([driver("transpose_f")], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    assert "mismatches 0" in r.stdout


@pytest.mark.parametrize("shape", [(6, 4, 5), (33, 17, 9)])
def test_host_transposes_match_oracle(oracle, shape):
    import mhdgpu
    u0, b0 = random_state(*shape, seed=1300 + shape[0])
    for which in (12, 13):
        ut, bt = mhdgpu.transpose_host(which, u0, b0, *shape)
        uo, bo = getattr(oracle, f"transpose{which}")(u0, b0, *shape)
        assert bitwise(ut, uo) and bitwise(bt, bo)
