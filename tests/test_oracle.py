"""Pins the oracle (oracle/mhd_oracle.c) before anything is compared with it.

* golden vectors written by the reference's own C functions (tests/golden/
  make_golden.py runs oracle/_ref, the reference's Heterogeneous_MHD/C
  sources compiled in place) -> bit-exact;
* the reference build itself, where present -> bit-exact on fresh seeds and
  shapes;
* mathematics: transposes are involutive permutations, uniform states are
  fixed points, fluidx conserves mass/momentum/energy, advectbyzx keeps the
  discrete div B and leaves u alone, a field at rest is left unchanged by
  advectbyzx (zero tvdb fluxes), the CFL speed has its closed form, the
  linear Alfven wave travels at v_A = 1 with second-order convergence, and the
  FP64 instantiation agrees with FP32 to rounding.
"""
import numpy as np
import pytest

from conftest import bitwise, same
from golden.states import random_state
import oracle_lib

F32 = np.float32


# ---------------------------------------------------------------- golden ----

def test_golden_fixture16_single_ops(oracle, golden, fixture16):
    u0, b0 = fixture16
    assert bitwise(u0, golden["fix16_u0"]) and bitwise(b0, golden["fix16_b0"])
    u, b = u0.copy(), b0.copy()
    oracle.fluidx(u, b, 16, 16, 16, F32(0.3))
    assert bitwise(u, golden["fix16_fluidx_u"])
    u, b = u0.copy(), b0.copy()
    oracle.advectbyzx(u, b, 16, 16, 16, F32(0.3))
    assert bitwise(b, golden["fix16_advect_b"])
    assert F32(oracle.calcfl(u0, b0, 16, 16, 16)) == golden["fix16_calcfl"]


def test_golden_fixture16_ten_steps(oracle, golden, fixture16):
    u, b = (a.copy() for a in fixture16)
    dts = oracle.run(u, b, 16, 16, 16, 10)
    assert bitwise(dts, golden["fix16_run10_dt"])
    assert bitwise(u, golden["fix16_run10_u"]) and bitwise(b, golden["fix16_run10_b"])


def test_golden_random_noncubic(oracle, golden):
    nx, ny, nz = 12, 8, 16
    u0, b0 = random_state(nx, ny, nz, seed=11, amp=0.3)
    assert bitwise(u0, golden["rnd_12x8x16_u0"]) and bitwise(b0, golden["rnd_12x8x16_b0"])
    u, b = u0.copy(), b0.copy()
    oracle.fluidx(u, b, nx, ny, nz, F32(0.25))
    assert bitwise(u, golden["rnd_12x8x16_fluidx_u"])
    u, b = u0.copy(), b0.copy()
    oracle.advectbyzx(u, b, nx, ny, nz, F32(0.25))
    assert bitwise(b, golden["rnd_12x8x16_advect_b"])
    ut, bt = oracle.transpose12(u0, b0, nx, ny, nz)
    assert bitwise(ut, golden["rnd_12x8x16_t12_u"]) and bitwise(bt, golden["rnd_12x8x16_t12_b"])
    assert F32(oracle.calcfl(u0, b0, nx, ny, nz)) == golden["rnd_12x8x16_calcfl"]


def test_golden_random_steps(oracle, golden):
    u0, b0 = golden["rnd8_u0"], golden["rnd8_b0"]
    u, b = u0.copy(), b0.copy()
    dts = oracle.run(u, b, 8, 8, 8, 5)
    assert bitwise(dts, golden["rnd8_run5_dt"])
    assert bitwise(u, golden["rnd8_run5_u"]) and bitwise(b, golden["rnd8_run5_b"])
    ut, bt = oracle.transpose13(u0, b0, 8, 8, 8)
    assert bitwise(ut, golden["rnd8_t13_u"]) and bitwise(bt, golden["rnd8_t13_b"])


# ------------------------------------------------------- reference build ----

needs_ref = pytest.mark.skipif(not oracle_lib.reference_available(),
                               reason="oracle/_ref not built (reference tree absent)")


@needs_ref
@pytest.mark.parametrize("shape,seed,amp", [((16, 16, 16), 21, 0.3), ((20, 6, 10), 22, 0.45),
                                            ((9, 5, 3), 23, 0.2), ((24, 24, 24), 24, 0.6),
                                            # pencil lengths of the generic GPU sweeps, up to 1024
                                            ((67, 5, 3), 25, 0.45), ((131, 4, 3), 26, 0.45),
                                            ((600, 3, 2), 27, 0.4), ((1024, 2, 2), 28, 0.4)])
def test_oracle_equals_reference_build_ops(oracle, shape, seed, amp):
    ref = oracle_lib.Reference()
    nx, ny, nz = shape
    u0, b0 = random_state(nx, ny, nz, seed=seed, amp=amp)
    for op, dt in (("fluidx", F32(0.21)), ("advectbyzx", F32(0.37))):
        ua, ba, ub, bb = u0.copy(), b0.copy(), u0.copy(), b0.copy()
        getattr(oracle, op)(ua, ba, nx, ny, nz, dt)
        getattr(ref, op)(ub, bb, nx, ny, nz, dt)
        assert bitwise(ua, ub) and bitwise(ba, bb), op
    assert oracle.calcfl(u0, b0, nx, ny, nz) == ref.calcfl(u0, b0, nx, ny, nz)
    ua, ba = u0.copy(), b0.copy()
    uo, bo = oracle.transpose12(ua, ba, nx, ny, nz)
    ref.transpose12(ua, ba, nx, ny, nz)
    assert bitwise(ua, uo) and bitwise(ba, bo)


@needs_ref
@pytest.mark.parametrize("n,seed,amp,steps", [(16, 31, 0.3, 4), (12, 32, 0.55, 6)])
def test_oracle_equals_reference_build_steps(oracle, n, seed, amp, steps):
    ref = oracle_lib.Reference()
    u0, b0 = random_state(n, n, n, seed=seed, amp=amp)
    ua, ba, ub, bb = u0.copy(), b0.copy(), u0.copy(), b0.copy()
    da = oracle.run(ua, ba, n, n, n, steps, tf=F32(3.0))  # the clamp engages on the last steps
    db = ref.run(ub, bb, n, n, n, steps, tf=F32(3.0))
    assert bitwise(da, db)
    assert bitwise(ua, ub) and bitwise(ba, bb)


@needs_ref
def test_oracle_equals_reference_build_fixture32(oracle, fixture32):
    ref = oracle_lib.Reference()
    ua, ba = (a.copy() for a in fixture32)
    ub, bb = (a.copy() for a in fixture32)
    assert bitwise(oracle.run(ua, ba, 32, 32, 32, 3), ref.run(ub, bb, 32, 32, 32, 3))
    assert bitwise(ua, ub) and bitwise(ba, bb)


# ------------------------------------------------------------ invariants ----

def test_transposes_are_involutions(oracle):
    for shape in ((6, 4, 5), (8, 8, 8)):
        u0, b0 = random_state(*shape, seed=41)
        nx, ny, nz = shape
        u1, b1 = oracle.transpose12(u0, b0, nx, ny, nz)
        u2, b2 = oracle.transpose12(u1, b1, ny, nx, nz)
        assert bitwise(u2, u0) and bitwise(b2, b0)
        u1, b1 = oracle.transpose13(u0, b0, nx, ny, nz)
        u2, b2 = oracle.transpose13(u1, b1, nz, ny, nx)
        assert bitwise(u2, u0) and bitwise(b2, b0)
        # component maps (transpose12.c:22-29): u (1,3,2,4,5), b (2,1,3)
        U0 = u0.reshape(nz, ny, nx, 5)
        U1 = oracle.transpose12(u0, b0, nx, ny, nz)[0].reshape(nz, nx, ny, 5)
        assert np.array_equal(U1[2, 3, 1], U0[2, 1, 3][[0, 2, 1, 3, 4]])


def test_uniform_state_is_fixed_point(oracle):
    n = 8
    u = np.tile(np.array([1.3, 0.2, -0.1, 0.05, 2.0], F32), n ** 3)
    b = np.tile(np.array([0.7, -0.4, 0.3], F32), n ** 3)
    u1, b1 = u.copy(), b.copy()
    oracle.fluidx(u1, b1, n, n, n, F32(0.2))
    oracle.advectbyzx(u1, b1, n, n, n, F32(0.2))
    assert same(u1, u) and same(b1, b)
    oracle.run(u1, b1, n, n, n, 2)
    assert same(u1, u) and same(b1, b)


def test_fluidx_conserves_and_advect_keeps_divb(oracle):
    nx, ny, nz = 16, 12, 10
    u0, b0 = random_state(nx, ny, nz, seed=51, amp=0.4)
    u, b = u0.copy(), b0.copy()
    m0 = oracle.monitors(u, b, nx, ny, nz)
    oracle.fluidx(u, b, nx, ny, nz, F32(0.2))
    m1 = oracle.monitors(u, b, nx, ny, nz)
    assert np.allclose(m1[:5], m0[:5], rtol=1e-6, atol=1e-5)
    assert bitwise(b, b0)
    # div B of the random state is not zero; advectbyzx must keep it cell by cell
    def div(bb):
        B = bb.reshape(nz, ny, nx, 3).astype(np.float64)
        return (np.roll(B[..., 0], -1, 2) - B[..., 0]) + (np.roll(B[..., 1], -1, 1) - B[..., 1]) + \
            (np.roll(B[..., 2], -1, 0) - B[..., 2])
    d0 = div(b)
    u2 = u.copy()
    oracle.advectbyzx(u, b, nx, ny, nz, F32(0.2))
    assert bitwise(u, u2)
    assert np.max(np.abs(div(b) - d0)) < 2e-6


def test_advect_at_rest_leaves_the_field(oracle):
    # tvdb at vg = 0 (tvdb.c:10-91): the upwind product, the half-step field
    # and the limited flux are all exactly zero, so every face field and the
    # EMF correction of b1 come out unchanged bit for bit, whatever b is
    nx, ny, nz = 12, 10, 9
    u, b = random_state(nx, ny, nz, seed=61, amp=0.5)
    U = u.reshape(-1, 5)
    U[:, 1:4] = 0.0
    u0, b0 = u.copy(), b.copy()
    oracle.advectbyzx(u, b, nx, ny, nz, F32(0.37))
    assert bitwise(u, u0) and bitwise(b, b0)
    # the check can fail: a small v_x (the sweep's own velocity, u2/u1)
    # moves the field
    U[:, 1] = F32(1e-3)
    oracle.advectbyzx(u, b, nx, ny, nz, F32(0.37))
    assert not bitwise(b, b0)


def test_cfl_closed_form(oracle):
    n = 4
    rho, v, B, p = 1.7, np.array([0.3, -0.8, 0.1]), np.array([0.6, 0.2, -0.5]), 0.9
    u = np.tile(np.r_[rho, rho * v, p / (2 / 3) + 0.5 * rho * v @ v + 0.5 * B @ B].astype(F32), n ** 3)
    b = np.tile(B.astype(F32), n ** 3)
    c = np.abs(v).max() + np.sqrt((2 * B @ B + 5 / 3 * p) / rho)
    assert abs(oracle.calcfl(u, b, n, n, n) - 1 / c) < 1e-6 / c
    # the cell-averaged B^2 variant agrees for a uniform field
    assert abs(oracle.calcfl(u, b, n, n, n, avg_b=True) - 1 / c) < 1e-6 / c


def test_linear_alfven_wave_speed_and_convergence(oracle):
    """dv_y = eps sin(2 pi (x - t)/L): v_A = B0/sqrt(rho0) = 1 along +x."""
    import ctypes
    from mhdgpu.ics import alfven_linear
    errs = {}
    for n in (16, 32):
        u, b = alfven_linear(n, eps=1e-4, ny=2, nz=2)
        t = ctypes.c_float(0.0)
        t_end = F32(n / 2.0)  # the wave crosses half the box
        step = oracle.lib.mhd_oracle_step_f32
        while t.value < t_end:
            step(u, b, n, 2, 2, ctypes.byref(t), t_end, 0)
        vy = (u.reshape(2, 2, n, 5)[0, 0, :, 2] / u.reshape(2, 2, n, 5)[0, 0, :, 0]).astype(np.float64)
        x = np.arange(n) + 1.0
        exact = 1e-4 * np.sin(2 * np.pi * (x - t.value) / n)
        errs[n] = np.sqrt(np.mean((vy - exact) ** 2)) / 1e-4
        # phase: the wave's centroid moved by t (to a small fraction of a cell)
        ph = np.angle(np.sum(vy * np.exp(-2j * np.pi * x / n)))
        ph0 = np.angle(np.sum(exact * np.exp(-2j * np.pi * x / n)))
        assert abs((ph - ph0 + np.pi) % (2 * np.pi) - np.pi) * n / (2 * np.pi) < 0.05
    assert errs[32] < 0.02
    # close to second order (the TVD limiter clips the extrema): error ratio
    # ~3.2 per halving of the cell size (4 would be exactly second order)
    assert errs[16] / errs[32] > 2.8


def test_fp64_instantiation_agrees_to_rounding(oracle):
    n = 12
    u0, b0 = random_state(n, n, n, seed=61, amp=0.3)
    u32, b32 = u0.copy(), b0.copy()
    u64, b64 = u0.astype(np.float64), b0.astype(np.float64)
    oracle.run(u32, b32, n, n, n, 3)
    oracle.run(u64, b64, n, n, n, 3)
    rel = np.linalg.norm(u32 - u64) / np.linalg.norm(u64)
    assert 0 < rel < 1e-5
    assert np.linalg.norm(b32 - b64) / np.linalg.norm(b64) < 1e-5


def test_fixture_reader_and_generator(oracle):
    from mhdgpu.ics import fixture_alfven
    import os
    from conftest import GOLDEN
    for n in (16, 32):
        rec = np.fromfile(os.path.join(GOLDEN, f"source_init_alfvenlinear_{n}cube.dat"), dtype="<f4")
        gen = fixture_alfven(n)
        assert np.max(np.abs(rec - gen)) < 1e-6
        u, b = oracle.read_ref_order(rec, n)
        U = u.reshape(n, n, n, 5)
        # the reference reader puts the file's fastest axis on k (init.c:33-54)
        assert np.ptp(U[:, :, :, 2], axis=(1, 2)).max() == 0 and np.ptp(U[:, :, :, 2], axis=0).max() > 0.1
