"""Writes tests/golden/golden_ref.npz: input/output vectors produced by the
reference's own C timestep functions (Heterogeneous_MHD/C, compiled from the
reference tree by oracle/Makefile into oracle/_ref/libmhd_ref.so).

Run in the container that holds the reference:
    make -C oracle && python tests/golden/make_golden.py

The vectors pin the oracle (and through it the GPU path) where the
reference tree is absent.  Inputs are the reference's fixture
data/source_init_alfvenlinear_16cube.dat (copied to tests/golden/) and
seeded random states written into the file themselves.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
from oracle_lib import Oracle, Reference  # noqa: E402
from golden.states import random_state  # noqa: E402
from golden.special_states import CFL_CASES, special_state  # noqa: E402


def main():
    O, Rf = Oracle(), Reference()
    out = {}
    rec = np.fromfile(os.path.join(HERE, "source_init_alfvenlinear_16cube.dat"), dtype="<f4")
    u0, b0 = O.read_ref_order(rec, 16)
    out["fix16_u0"], out["fix16_b0"] = u0, b0
    ua, ba = u0.copy(), b0.copy()
    Rf.fluidx(ua, ba, 16, 16, 16, np.float32(0.3))
    out["fix16_fluidx_u"] = ua
    ua, ba = u0.copy(), b0.copy()
    Rf.advectbyzx(ua, ba, 16, 16, 16, np.float32(0.3))
    out["fix16_advect_b"] = ba
    out["fix16_calcfl"] = np.float32(Rf.calcfl(u0, b0, 16, 16, 16))
    ua, ba = u0.copy(), b0.copy()
    out["fix16_run10_dt"] = Rf.run(ua, ba, 16, 16, 16, 10)
    out["fix16_run10_u"], out["fix16_run10_b"] = ua, ba

    # non-cubic random state: single ops and the transposes
    nx, ny, nz = 12, 8, 16
    u0, b0 = random_state(nx, ny, nz, seed=11, amp=0.3)
    out["rnd_12x8x16_u0"], out["rnd_12x8x16_b0"] = u0, b0
    ua, ba = u0.copy(), b0.copy()
    Rf.fluidx(ua, ba, nx, ny, nz, np.float32(0.25))
    out["rnd_12x8x16_fluidx_u"] = ua
    ua, ba = u0.copy(), b0.copy()
    Rf.advectbyzx(ua, ba, nx, ny, nz, np.float32(0.25))
    out["rnd_12x8x16_advect_b"] = ba
    ua, ba = u0.copy(), b0.copy()
    Rf.transpose12(ua, ba, nx, ny, nz)
    out["rnd_12x8x16_t12_u"], out["rnd_12x8x16_t12_b"] = ua, ba
    out["rnd_12x8x16_calcfl"] = np.float32(Rf.calcfl(u0, b0, nx, ny, nz))

    # strong random state on a cube: whole steps (exercises the limiters)
    u0, b0 = random_state(8, 8, 8, seed=12, amp=0.5)
    out["rnd8_u0"], out["rnd8_b0"] = u0, b0
    ua, ba = u0.copy(), b0.copy()
    out["rnd8_run5_dt"] = Rf.run(ua, ba, 8, 8, 8, 5)
    out["rnd8_run5_u"], out["rnd8_run5_b"] = ua, ba
    ua, ba = u0.copy(), b0.copy()
    Rf.transpose13(ua, ba, 8, 8, 8)
    out["rnd8_t13_u"], out["rnd8_t13_b"] = ua, ba

    # non-positive densities, NaN and infinities: the CFL max's NaN semantics
    # (calcfl.c:55) and the ranges beyond the GPU's fast-division window
    for name in CFL_CASES:
        u0, b0 = special_state(name, 12, 6, 10, seed=5)
        out[f"spc_{name}_calcfl"] = np.float32(Rf.calcfl(u0, b0, 12, 6, 10))
    for name in ("neg_density", "neg_density_many", "tiny_density", "huge_momentum"):
        ua, ba = special_state(name, 10, 10, 10, seed=5)
        out[f"spc_{name}_run3_dt"] = Rf.run(ua, ba, 10, 10, 10, 3)
        out[f"spc_{name}_run3_u"], out[f"spc_{name}_run3_b"] = ua, ba

    np.savez_compressed(os.path.join(HERE, "golden_ref.npz"), **out)
    print("wrote", sorted(out))


if __name__ == "__main__":
    main()
