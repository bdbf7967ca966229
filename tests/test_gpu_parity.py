"""GPU parity: every op of the timestep through the C ABI (libmhdgpu.so)
against the oracle on the same seeded inputs.

The bar is the reference's own results: every value equal to the oracle's
(which equals the reference C path bit for bit), the sign of exact zeros
aside, and the dt sequence bit-exact.  Pencil lengths cover every cells-per-
lane variant (K = 1..8) and partial waves (L < 64 lanes).
"""
import math
import os

import numpy as np
import pytest

from conftest import GOLDEN, bitwise, same
from golden.states import random_state

pytestmark = pytest.mark.gpu
F32 = np.float32


@pytest.fixture(scope="module")
def M():
    import mhdgpu
    free, total = mhdgpu.mem_info()  # raises if the library or the GPU is missing
    assert total > 0
    return mhdgpu


def dscalar(M, x, dtype=F32):
    return M.DeviceArray.from_host(np.array([x], dtype))


def gpu_op(M, op, u0, b0, nx, ny, nz, dt):
    du, db = M.DeviceArray.from_host(u0), M.DeviceArray.from_host(b0)
    getattr(M, op)(du, db, nx, ny, nz, dscalar(M, dt))
    return du.to_host(), db.to_host()


def test_rcp_f64_bound_of_the_limiter_quotient(M):
    # the FP32 limiter forms 2pm/(p+m) from the FP64 reciprocal instruction
    # and one Newton step; its equality with the IEEE quotient needs the
    # instruction within 2^-22 of 1/D on every float-valued D (exhaustive)
    err = M.selfcheck()
    print("rcp_f64 max relative error 2^%.2f" % math.log2(err), flush=True)
    assert 0.0 < err < 2.0 ** -22


def test_shared_divisor_quotient_is_ieee(M):
    # the FP32 divisions of mhdflux and of the flux normalisation take one
    # residual correction from the refined reciprocal; that equals the IEEE
    # quotient for every dividend mantissa against every 61st divisor
    # mantissa here (all 2^46 pairs with stride 1, bench.py --mode selfcheck),
    # and the reciprocal instruction scales exactly, which extends it to the
    # exponent ranges the sweeps admit
    bad_q, bad_rcp = M.selfcheck_division(61)
    print("one-correction quotient mismatches %d, rcp equivariance mismatches %d" % (bad_q, bad_rcp), flush=True)
    assert bad_q == 0 and bad_rcp == 0


def test_freezing_speed_sqrt_is_ieee(M):
    # the root in mhdflux's freezing speed (mhdflux.c:45) takes a float-valued
    # argument, and its shortened sequence must give the IEEE double root on
    # every positive finite float
    bad = M.selfcheck_sqrt()
    print("shortened sqrt mismatches over all positive floats: %d" % bad, flush=True)
    assert bad == 0


def limiter_pairs(rng):
    """FP32 operand pairs of the van Leer limiter: random bit patterns (every
    exponent, signs, subnormals, infinities, NaN), sums that overflow FP32,
    products on either side of the 2^-150 underflow tie, exact quotients and
    near-equal operands"""
    bits = rng.integers(0, 1 << 32, size=(200000, 2), dtype=np.uint64).astype(np.uint32)
    p, m = list(bits[:, 0].view(np.float32)), list(bits[:, 1].view(np.float32))
    big = rng.uniform(1.0, 2.0, size=(5000, 2)) * 2.0 ** 127
    sgn = np.where(rng.random(5000) < 0.5, -1.0, 1.0)
    p += list(np.float32(big[:, 0] * sgn))
    m += list(np.float32(big[:, 1] * sgn))
    for x in rng.random(5000):
        a = np.float32(np.ldexp(1 + x, -75))
        b = np.float32(2.0 ** -150 / float(a))
        for d in (-1, 0, 1):
            p.append(a)
            m.append(np.float32(b * np.float32(1 + d * 2.0 ** -23)))
    near = rng.uniform(0.5, 2.0, 5000).astype(np.float32)
    p += list(near)
    m += list(np.float32(near * np.float32(1 + 2.0 ** -22)))
    special = [np.inf, -np.inf, np.nan, 0.0, -0.0, 2.0 ** -149, 3.4028235e38, 1.0, 3.0, 0.75]
    for a in special:
        for b in special:
            p.append(np.float32(a))
            m.append(np.float32(b))
    return np.array(p, np.float32), np.array(m, np.float32)


def test_limiter_forms_equal_the_reference_expression(M):
    # tvd1.c:71-81, tvdb.c:72-76: d = 2.0*p*m/(p+m) in double with p+m summed
    # in FP32, where the FP32 p*m > 0, else 0; both device forms (the test
    # read off the exact FP64 numerator, the refined reciprocal, the IEEE
    # division for an infinite sum, the clamp of the branch-free form) must
    # give it bit for bit, NaN where it is NaN
    p, m = limiter_pairs(np.random.default_rng(23))
    with np.errstate(all="ignore"):
        taken = (p * m) > np.float32(0)
        q = (2.0 * p.astype(np.float64) * m.astype(np.float64) / (p + m).astype(np.float64)).astype(np.float32)
    want = np.where(taken, q, np.float32(0))
    out, out_nb = M.selfcheck_limiter(p, m)
    print("limiter pairs", p.size, "taken", int(taken.sum()), flush=True)
    for got in (out, out_nb):
        nan = np.isnan(want)
        assert np.array_equal(np.isnan(got), nan)
        assert np.array_equal(got[~nan].view(np.uint32), want[~nan].view(np.uint32))


SHAPES = [(16, 4, 3), (48, 5, 4), (64, 3, 3), (96, 3, 5), (128, 2, 3), (192, 2, 2), (256, 3, 2), (320, 1, 2),
          (384, 2, 1), (512, 1, 2), (24, 1, 1), (40, 7, 1)]


@pytest.mark.parametrize("shape", SHAPES)
def test_fluidx_matches_oracle(M, oracle, shape):
    nx, ny, nz = shape
    print("fluidx", shape, flush=True)
    u0, b0 = random_state(nx, ny, nz, seed=100 + nx, amp=0.4)
    dt = F32(0.23)
    ug, bg = gpu_op(M, "fluidx", u0, b0, nx, ny, nz, dt)
    uo, bo = u0.copy(), b0.copy()
    oracle.fluidx(uo, bo, nx, ny, nz, dt)
    assert same(ug, uo)
    assert bitwise(bg, b0)


@pytest.mark.parametrize("shape", SHAPES)
def test_advectbyzx_matches_oracle(M, oracle, shape):
    nx, ny, nz = shape
    print("advectbyzx", shape, flush=True)
    u0, b0 = random_state(nx, ny, nz, seed=200 + nx, amp=0.5)
    dt = F32(0.31)
    ug, bg = gpu_op(M, "advectbyzx", u0, b0, nx, ny, nz, dt)
    uo, bo = u0.copy(), b0.copy()
    oracle.advectbyzx(uo, bo, nx, ny, nz, dt)
    assert bitwise(ug, u0)
    assert same(bg, bo)


@pytest.mark.parametrize("shape", [(64, 40, 9), (256, 3, 70)])
def test_advectbyzx_at_rest_leaves_the_field(M, shape):
    # zero velocity: every tvdb flux is exactly zero, so b must come back bit
    # for bit (the x-frame tiles and the y/z column streams alike)
    nx, ny, nz = shape
    print("advectbyzx at rest", shape, flush=True)
    u0, b0 = random_state(nx, ny, nz, seed=300 + nx, amp=0.5)
    u0.reshape(-1, 5)[:, 1:4] = 0.0
    ug, bg = gpu_op(M, "advectbyzx", u0, b0, nx, ny, nz, F32(0.37))
    assert bitwise(ug, u0) and bitwise(bg, b0)


def test_single_ops_on_reference_fixtures(M, oracle, golden, fixture16, fixture32):
    u0, b0 = fixture16
    ug, _ = gpu_op(M, "fluidx", u0, b0, 16, 16, 16, F32(0.3))
    assert same(ug, golden["fix16_fluidx_u"])
    _, bg = gpu_op(M, "advectbyzx", u0, b0, 16, 16, 16, F32(0.3))
    assert same(bg, golden["fix16_advect_b"])
    u0, b0 = fixture32
    ug, _ = gpu_op(M, "fluidx", u0, b0, 32, 32, 32, F32(0.41))
    uo = u0.copy()
    oracle.fluidx(uo, b0.copy(), 32, 32, 32, F32(0.41))
    assert same(ug, uo)


# square planes (nx == ny for transpose12, nx == nz for transpose13) take the
# in-place pair swap, with 16-byte rows where nx is a multiple of 32 (32^3,
# 96^3, 64x32x64) and value by value elsewhere (40x40x7, 20x9x20); the other
# shapes transpose out of place and copy back
@pytest.mark.parametrize("shape", [(6, 5, 4), (32, 32, 32), (33, 17, 9), (40, 3, 70), (96, 96, 96), (64, 32, 64),
                                   (40, 40, 7), (20, 9, 20)])
def test_transposes_match_oracle(M, oracle, shape):
    nx, ny, nz = shape
    u0, b0 = random_state(nx, ny, nz, seed=300)
    for which in (12, 13):
        print("transpose", which, shape, flush=True)
        du, db = M.DeviceArray.from_host(u0), M.DeviceArray.from_host(b0)
        ut, bt = M.DeviceArray(u0.nbytes), M.DeviceArray(b0.nbytes)
        getattr(M, f"transpose{which}")(ut, bt, du, db, nx, ny, nz)
        uo, bo = getattr(oracle, f"transpose{which}")(u0, b0, nx, ny, nz)
        assert bitwise(du.to_host(), uo) and bitwise(db.to_host(), bo)
        assert bitwise(ut.to_host(), uo) and bitwise(bt.to_host(), bo)


@pytest.mark.parametrize("shape,avg", [((16, 16, 16), False), ((33, 8, 5), False), ((24, 12, 6), True),
                                       ((128, 64, 8), False)])
def test_cfl_matches_oracle(M, oracle, shape, avg):
    nx, ny, nz = shape
    u0, b0 = random_state(nx, ny, nz, seed=400, amp=0.5)
    d = dscalar(M, 0.0)
    c = M.cfl(M.DeviceArray.from_host(u0), M.DeviceArray.from_host(b0), nx, ny, nz, d, avg_b=avg)
    assert F32(c) == F32(oracle.calcfl(u0, b0, nx, ny, nz, avg_b=avg))
    assert d.to_host()[0] == F32(c)


def test_ctx_ten_steps_equal_reference_golden(M, golden, fixture16):
    sim = M.Simulation(16, 16, 16)
    sim.set_state(*fixture16)
    sim.step(10)
    u, b = sim.get_state()
    assert bitwise(sim.dts(), golden["fix16_run10_dt"])
    assert same(u, golden["fix16_run10_u"]) and same(b, golden["fix16_run10_b"])
    sim.close()


def test_config1_alfven_32cube_200_steps(M, oracle):
    """The CPU-runnable workload: 32^3 linear Alfven wave, 200 steps, FP32."""
    from mhdgpu.ics import alfven_linear
    u0, b0 = alfven_linear(32)
    sim = M.Simulation(32, 32, 32)
    sim.set_state(u0, b0)
    sim.step(200)
    ug, bg = sim.get_state()
    uo, bo = u0.copy(), b0.copy()
    do = oracle.run(uo, bo, 32, 32, 32, 200)
    assert bitwise(sim.dts(), do)
    assert same(ug, uo) and same(bg, bo)
    rel = np.linalg.norm(ug.astype(np.float64) - uo) / np.linalg.norm(uo)
    assert rel <= 1e-5
    sim.close()


@pytest.mark.parametrize("shape,seed,amp,steps,tf", [((24, 16, 32), 501, 0.4, 5, math.inf),
                                                     ((16, 16, 16), 502, 0.55, 6, 3.0),
                                                     ((48, 96, 64), 503, 0.3, 2, math.inf),
                                                     ((42, 22, 26), 504, 0.45, 3, math.inf),
                                                     # pencils of 2 and 3 cells; a partial 64-pencil tile
                                                     ((70, 2, 3), 505, 0.4, 3, math.inf),
                                                     # an odd plane count
                                                     ((64, 100, 45), 506, 0.35, 2, math.inf),
                                                     # y/z fluid sweeps of pencils that fill the wave
                                                     # (64K cells) over pairs of tiles: odd tile counts
                                                     # (the last workgroup has one), K = 1, 4, 2
                                                     ((24, 64, 3), 507, 0.4, 3, math.inf),
                                                     ((8, 256, 5), 508, 0.4, 2, math.inf),
                                                     ((56, 128, 64), 509, 0.3, 2, math.inf)])
def test_ctx_steps_match_oracle(M, oracle, shape, seed, amp, steps, tf):
    nx, ny, nz = shape
    u0, b0 = random_state(nx, ny, nz, seed=seed, amp=amp)
    sim = M.Simulation(nx, ny, nz)
    sim.set_state(u0, b0)
    sim.reset_clock(tf)
    sim.step(steps)
    ug, bg = sim.get_state()
    uo, bo = u0.copy(), b0.copy()
    do = oracle.run(uo, bo, nx, ny, nz, steps, tf=F32(tf))
    assert bitwise(sim.dts(), do)
    assert same(ug, uo) and same(bg, bo)
    t, t64 = sim.time()
    assert abs(t64 - 2 * do.astype(np.float64).sum()) < 1e-12
    sim.close()


_FORM_ORACLE = {}


# Every kernel form of the y/z advectbyzx, forced at small sizes (the automatic
# choice streams only boxes of >= 2048 pencil segments): one workgroup wave
# (nx 64, 70), three and five waves with a partial last one (nx 130, 300: the
# fused stream's 4- and 16-wave builds), odd plane counts, pencils of 2 cells,
# and pencils cut into segments with a short last one (ny 100).
@pytest.mark.parametrize("form", ["tiles", "two_pass", "fused"])
@pytest.mark.parametrize("shape", [(64, 100, 45), (130, 40, 33), (300, 36, 20), (70, 2, 3)])
def test_advect_forms_match_oracle(M, oracle, shape, form):
    nx, ny, nz = shape
    print("advect form", form, shape, flush=True)
    u0, b0 = random_state(nx, ny, nz, seed=600 + nx, amp=0.4)
    if shape not in _FORM_ORACLE:
        uo, bo = u0.copy(), b0.copy()
        do = oracle.run(uo, bo, nx, ny, nz, 2)
        _FORM_ORACLE[shape] = (uo, bo, do)
    uo, bo, do = _FORM_ORACLE[shape]
    sim = M.Simulation(nx, ny, nz)
    sim.set_advect_form(form)
    sim.set_state(u0, b0)
    sim.step(2)
    ug, bg = sim.get_state()
    dg = sim.dts()
    sim.close()
    assert bitwise(dg, do)
    assert same(ug, uo) and same(bg, bo)


@pytest.mark.parametrize("avg", [False, True])
def test_set_state_between_steps_recomputes_the_cfl(M, avg):
    # the CFL of a step comes from the previous step's last sweep; a state set
    # in between must not inherit it: the run continues exactly as a fresh
    # context started from that state
    shape = (24, 20, 16)
    ua, ba = random_state(*shape, seed=701, amp=0.5)
    ub, bb = random_state(*shape, seed=702, amp=0.05)  # slower state: a larger dt
    sim = M.Simulation(*shape, avg_b=avg)
    sim.set_state(ua, ba)
    sim.step(3)
    sim.set_state(ub, bb)  # no reset_clock: set_state alone must recompute the CFL
    sim.step(2)
    u1, b1 = sim.get_state()
    d1 = sim.dts()[3:]
    sim.close()
    fresh = M.Simulation(*shape, avg_b=avg)
    fresh.set_state(ub, bb)
    fresh.step(2)
    u2, b2 = fresh.get_state()
    d2 = fresh.dts()
    fresh.close()
    assert bitwise(d1, d2) and bitwise(u1, u2) and bitwise(b1, b2)


def test_ctx_fp64_matches_fp64_oracle(M, oracle):
    n = 12
    u0, b0 = random_state(n, n, n, seed=601, amp=0.3)
    sim = M.Simulation(n, n, n, precision=8)
    sim.set_state(u0.astype(np.float64), b0.astype(np.float64))
    sim.step(3)
    ug, bg = sim.get_state()
    uo, bo = u0.astype(np.float64), b0.astype(np.float64)
    do = oracle.run(uo, bo, n, n, n, 3)
    assert bitwise(sim.dts(), do)
    assert same(ug, uo) and same(bg, bo)
    sim.close()


_FORM_ORACLE64 = {}


# The FP64 state through the vector paths and every y/z advectbyzx form: at
# (128, 64, 10) the x pencils of 128 cells take K = 2 with 16-byte loads and
# the y pencils of 64 cells the full-wave two-vector tiles; at (130, 40, 33)
# the fused stream's partial third wave and odd plane counts
@pytest.mark.parametrize("form", ["tiles", "two_pass", "fused"])
@pytest.mark.parametrize("shape", [(128, 64, 10), (130, 40, 33)])
def test_fp64_forms_match_fp64_oracle(M, oracle, shape, form):
    nx, ny, nz = shape
    print("fp64 advect form", form, shape, flush=True)
    u0, b0 = random_state(nx, ny, nz, seed=610 + nx, amp=0.35)
    u0, b0 = u0.astype(np.float64), b0.astype(np.float64)
    if shape not in _FORM_ORACLE64:
        uo, bo = u0.copy(), b0.copy()
        do = oracle.run(uo, bo, nx, ny, nz, 2)
        _FORM_ORACLE64[shape] = (uo, bo, do)
    uo, bo, do = _FORM_ORACLE64[shape]
    sim = M.Simulation(nx, ny, nz, precision=8)
    sim.set_advect_form(form)
    sim.set_state(u0, b0)
    sim.step(2)
    ug, bg = sim.get_state()
    dg = sim.dts()
    sim.close()
    assert bitwise(dg, do)
    assert same(ug, uo) and same(bg, bo)


def test_avg_b_cfl_variant_in_driver(M, oracle):
    n = 16
    u0, b0 = random_state(n, n, n, seed=602, amp=0.4)
    sim = M.Simulation(n, n, n, avg_b=True)
    sim.set_state(u0, b0)
    sim.step(3)
    uo, bo = u0.copy(), b0.copy()
    do = oracle.run(uo, bo, n, n, n, 3, avg_b=True)
    assert bitwise(sim.dts(), do)
    ug, bg = sim.get_state()
    assert same(ug, uo) and same(bg, bo)
    sim.close()


def test_main_drop_in(M, oracle, fixture16):
    u, b = (a.copy() for a in fixture16)
    dts = M.main(u, b, 16, 16, 16, 4)
    uo, bo = (a.copy() for a in fixture16)
    do = oracle.run(uo, bo, 16, 16, 16, 4)
    assert bitwise(dts, do) and same(u, uo) and same(b, bo)
    # run-to-tf mode of the reference loop (t < tf)
    u, b = (a.copy() for a in fixture16)
    dts = M.main(u, b, 16, 16, 16, 0, tf=5.0)
    uo, bo = (a.copy() for a in fixture16)
    do = oracle.run(uo, bo, 16, 16, 16, len(dts), tf=F32(5.0))
    assert bitwise(dts, do) and same(u, uo)


def test_monitors_and_determinism(M, oracle):
    n = 32
    u0, b0 = random_state(n, n, n, seed=701, amp=0.3)
    outs = []
    for _ in range(2):
        sim = M.Simulation(n, n, n)
        sim.set_state(u0, b0)
        m0 = sim.monitors()
        sim.step(3)
        outs.append(sim.get_state())
        m1 = sim.monitors()
        sim.close()
    assert bitwise(outs[0][0], outs[1][0]) and bitwise(outs[0][1], outs[1][1])
    mo = oracle.monitors(u0, b0, n, n, n)
    assert np.allclose(m0, mo, rtol=1e-12, atol=1e-12)
    assert abs(m1[0] - m0[0]) <= 1e-6 * abs(m0[0])  # mass
    assert abs(m1[1] - m0[1]) <= 1e-5 * abs(m0[1])  # energy
    assert abs(m1[5] - m0[5]) < 1e-5  # the CT update keeps div B


def test_full_size_256_step_properties_and_parity(M, oracle):
    """The 256^3 workload (random-phase velocity, curl-A field): one step on
    the GPU against the oracle, plus the size-independent checks."""
    from mhdgpu.ics import random_phase
    n = 256
    u0, b0 = random_phase(n, seed=2)
    sim = M.Simulation(n, n, n)
    sim.set_state(u0, b0)
    m0 = sim.monitors()
    sim.step(1)
    ug, bg = sim.get_state()
    m1 = sim.monitors()
    dg = sim.dts()
    sim.close()
    uo, bo = u0.copy(), b0.copy()
    do = oracle.run(uo, bo, n, n, n, 1)
    assert bitwise(dg, do)
    assert same(ug, uo) and same(bg, bo)
    assert m0[5] < 1e-5 and m1[5] < 1e-5  # div B from curl A stays at rounding level
    assert abs(m1[0] - m0[0]) <= 1e-6 * m0[0]


@pytest.mark.timeout(900)
def test_max_size_512_step_matches_oracle(M, oracle):
    """The precision-stress workload at its full size: 512^3 decaying
    turbulence (log-normal density, sigma 0.5; solenoidal velocity at rms
    Mach 1; field from a random vector potential at beta 1), one step
    (pencils of 512 cells, K = 8; 5 x 512^3 = 6.7e8 values, int32 indices)
    against the oracle, bit for bit, with mass conserved and div B kept."""
    from mhdgpu.ics import turbulence
    n = 512
    u0, b0 = turbulence(n, seed=4, progress=lambda m: print("512^3:", m, flush=True))
    sim = M.Simulation(n, n, n)
    sim.set_state(u0, b0)
    m0 = sim.monitors()
    sim.step(1)
    ug, bg = sim.get_state()
    m1 = sim.monitors()
    dg = sim.dts()
    sim.close()
    print("512^3: oracle", flush=True)
    uo, bo = u0.copy(), b0.copy()
    del u0, b0
    do = oracle.run(uo, bo, n, n, n, 1)
    assert bitwise(dg, do)
    assert same(ug, uo) and same(bg, bo)
    assert abs(m1[0] - m0[0]) <= 1e-6 * m0[0]
    assert m1[5] < 1e-5  # the curl-A field's div B stays at the rounding level
