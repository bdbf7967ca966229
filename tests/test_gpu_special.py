"""GPU parity on states the ordinary tests never reach, through the C ABI,
against the oracle (which equals the reference build on every one of them,
tests/test_oracle_pins.py):

* non-positive, zero, infinite and NaN values: the CFL max's NaN semantics
  (Heterogeneous_MHD/C/calcfl.c:55, max_2num: a NaN speed resets the running
  max), in the step-level mhd_cfl, the context's captured step (standalone
  CFL after set_state, the CFL fused into the last sweep afterwards), the
  cell-averaged mode and the batched engine;
* densities and momenta far outside the window of the exact fast divisions
  (|u1| in [2^-30, 2^30), numerators below 2^60), so the waves take the
  true-division fallback of mhdflux and of the flux normalisation.

The bar: every value equal to the oracle's, NaN where the oracle has NaN.
"""
import numpy as np
import pytest

from conftest import same
from golden.special_states import CFL_CASES, EXTRA_CASES, special_state

pytestmark = pytest.mark.gpu
F32 = np.float32


@pytest.fixture(scope="module")
def M():
    import mhdgpu
    mhdgpu.mem_info()
    return mhdgpu


def f32_same(a, b):
    a, b = F32(a), F32(b)
    return a == b or (np.isnan(a) and np.isnan(b))


@pytest.mark.parametrize("name", CFL_CASES)
@pytest.mark.parametrize("shape,avg", [((12, 6, 10), False), ((33, 8, 5), True), ((64, 16, 12), False)])
def test_step_level_cfl_special_states(M, oracle, name, shape, avg):
    print("mhd_cfl", name, shape, avg, flush=True)
    u0, b0 = special_state(name, *shape, seed=5)
    d = M.DeviceArray.from_host(np.zeros(1, F32))
    c = M.cfl(M.DeviceArray.from_host(u0), M.DeviceArray.from_host(b0), *shape, d, avg_b=avg)
    assert f32_same(c, oracle.calcfl(u0, b0, *shape, avg_b=avg))
    assert f32_same(d.to_host()[0], c)


@pytest.mark.parametrize("name", CFL_CASES)
@pytest.mark.parametrize("shape,avg", [((10, 10, 10), False), ((24, 16, 32), False), ((16, 12, 8), True)])
def test_ctx_steps_special_states(M, oracle, name, shape, avg):
    print("ctx steps", name, shape, avg, flush=True)
    u0, b0 = special_state(name, *shape, seed=5)
    sim = M.Simulation(*shape, avg_b=avg)
    sim.set_state(u0, b0)
    sim.step(3)
    ug, bg = sim.get_state()
    dg = sim.dts()
    sim.close()
    uo, bo = u0.copy(), b0.copy()
    do = oracle.run(uo, bo, *shape, 3, avg_b=avg)
    assert same(dg, do)
    assert same(ug, uo) and same(bg, bo)


@pytest.mark.parametrize("name", ["tiny_density", "huge_momentum"])
@pytest.mark.parametrize("shape", [(64, 8, 6), (256, 3, 2), (24, 20, 9)])
def test_fast_division_fallbacks(M, oracle, name, shape):
    # single ops on the step-level path (every wave holding a tiny density or
    # a huge momentum recomputes with true divisions) ...
    print("fallback ops", name, shape, flush=True)
    u0, b0 = special_state(name, *shape, seed=7)
    dt = F32(1e-3)
    for op in ("fluidx", "advectbyzx"):
        du, db = M.DeviceArray.from_host(u0), M.DeviceArray.from_host(b0)
        getattr(M, op)(du, db, *shape, M.DeviceArray.from_host(np.array([dt], F32)))
        uo, bo = u0.copy(), b0.copy()
        getattr(oracle, op)(uo, bo, *shape, dt)
        assert same(du.to_host(), uo) and same(db.to_host(), bo), op
    # ... and whole steps of the device-resident path (the SoA sweeps in all
    # three frames), dt from the CFL of that state
    sim = M.Simulation(*shape)
    sim.set_state(u0, b0)
    sim.step(2)
    ug, bg = sim.get_state()
    dg = sim.dts()
    sim.close()
    uo, bo = u0.copy(), b0.copy()
    do = oracle.run(uo, bo, *shape, 2)
    assert same(dg, do) and same(ug, uo) and same(bg, bo)


@pytest.mark.parametrize("name", EXTRA_CASES + ["inf_density", "huge_momentum"])
@pytest.mark.parametrize("shape", [(64, 8, 6), (24, 20, 9)])
def test_overflow_states_single_ops(M, oracle, name, shape):
    # FP32 products, differences and limiter sums that overflow inside one
    # fluidx or advectbyzx: a single op leaves most pencils finite, so every
    # inf and NaN position is checked against the oracle
    print("overflow ops", name, shape, flush=True)
    u0, b0 = special_state(name, *shape, seed=7)
    dt = F32(0.2)
    for op in ("fluidx", "advectbyzx"):
        du, db = M.DeviceArray.from_host(u0), M.DeviceArray.from_host(b0)
        getattr(M, op)(du, db, *shape, M.DeviceArray.from_host(np.array([dt], F32)))
        uo, bo = u0.copy(), b0.copy()
        getattr(oracle, op)(uo, bo, *shape, dt)
        assert same(du.to_host(), uo) and same(db.to_host(), bo), op


@pytest.mark.parametrize("form", ["tiles", "two_pass", "fused"])
@pytest.mark.parametrize("name", EXTRA_CASES + ["inf_density", "nan_energy", "huge_momentum", "tiny_density"])
def test_special_states_every_advect_form(M, oracle, name, form):
    # the y/z advectbyzx forms, forced on a small box (the automatic choice
    # streams only large ones), on states with infinite, NaN, huge and tiny
    # values: two whole steps against the oracle
    shape = (64, 20, 9)
    print("advect form", form, name, flush=True)
    u0, b0 = special_state(name, *shape, seed=11)
    sim = M.Simulation(*shape)
    sim.set_advect_form(form)
    sim.set_state(u0, b0)
    sim.step(2)
    ug, bg = sim.get_state()
    dg = sim.dts()
    sim.close()
    uo, bo = u0.copy(), b0.copy()
    do = oracle.run(uo, bo, *shape, 2)
    assert same(dg, do) and same(ug, uo) and same(bg, bo)


def test_batch_box_with_nan_cfl(M, oracle):
    # one box of the batch has a negative density (a NaN CFL speed): its dt
    # follows the reference's sequential max, the other boxes are untouched
    dims = [(16, 16, 16), (10, 10, 10), (24, 16, 32), (32, 32, 32)]
    states = [special_state("neg_density" if i == 1 else "nan_energy" if i == 3 else "tiny_density", *d, seed=9 + i)
              for i, d in enumerate(dims)]
    bt = M.Batch(dims)
    for i, (u, b) in enumerate(states):
        bt.set_state(i, u, b)
    bt.step(3)
    bt.sync()
    for i, (d, (u0, b0)) in enumerate(zip(dims, states)):
        ug, bg = bt.get_state(i)
        uo, bo = u0.copy(), b0.copy()
        do = oracle.run(uo, bo, *d, 3)
        assert same(bt.dts(i), do), i
        assert same(ug, uo) and same(bg, bo), i
    bt.close()


def test_main_reports_nonfinite_dt(M, oracle):
    # mhd_main runs the reference's arithmetic on a state with a NaN energy,
    # writes the state and the dts back, and returns MHD_ERR_NONFINITE
    import ctypes
    n = 10
    u0, b0 = special_state("nan_energy", n, n, n, seed=5)
    u, b = u0.copy(), b0.copy()
    dts = np.zeros(3, F32)
    done = ctypes.c_int()
    ni = ctypes.c_int(n)
    rc = M.lib().mhd_main(u.ctypes.data, b.ctypes.data, ctypes.byref(ni), ctypes.byref(ni), ctypes.byref(ni), 3,
                          ctypes.c_float(np.inf), dts.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), 3,
                          ctypes.byref(done))
    assert rc == -8 and done.value == 3
    assert "dt" in M.lib().mhd_last_error().decode()
    uo, bo = u0.copy(), b0.copy()
    do = oracle.run(uo, bo, n, n, n, 3)
    assert same(dts, do) and same(u, uo) and same(b, bo)
    # a finite run returns MHD_OK through the same path
    u, b = special_state("neg_density", n, n, n, seed=5)
    assert np.all(np.isfinite(M.main(u, b, n, n, n, 3)))
