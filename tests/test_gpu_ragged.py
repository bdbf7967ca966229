"""GPU parity for every pencil length the reference accepts (2 <= n <= 1024,
one block of n threads in HIP/cu/cuda_fluidx.cu:15-17), beyond the fast
kernels' n = L*K with K <= 8: the generic sweeps hold K = 2, 4, 8 or 16 cells
per lane with the last lane ragged.  Step-level ops along x (AoS) and whole
steps of the device-resident context (SoA, every frame), against the oracle,
bit for bit (NaN equal to NaN, the sign of exact zeros aside).
"""
import numpy as np
import pytest

from conftest import same
from golden.states import random_state

pytestmark = pytest.mark.gpu
F32 = np.float32


@pytest.fixture(scope="module")
def M():
    import mhdgpu
    mhdgpu.mem_info()
    return mhdgpu


# K = 2 (67, 97: one valid cell in the last lane), 4 (131: three), 8 (263),
# 16 (600, 1000 ragged; 1024 whole)
@pytest.mark.parametrize("shape", [(67, 5, 3), (97, 3, 4), (131, 4, 3), (263, 3, 2), (600, 3, 2), (1000, 2, 3),
                                   (1024, 2, 2)])
def test_step_level_ops_any_pencil_length(M, oracle, shape):
    nx, ny, nz = shape
    print("ragged ops", shape, flush=True)
    u0, b0 = random_state(nx, ny, nz, seed=800 + nx, amp=0.45)
    for op, dt in (("fluidx", F32(0.23)), ("advectbyzx", F32(0.31))):
        du, db = M.DeviceArray.from_host(u0), M.DeviceArray.from_host(b0)
        getattr(M, op)(du, db, nx, ny, nz, M.DeviceArray.from_host(np.array([dt], F32)))
        uo, bo = u0.copy(), b0.copy()
        getattr(oracle, op)(uo, bo, nx, ny, nz, dt)
        assert same(du.to_host(), uo) and same(db.to_host(), bo), op


# ragged lengths in each frame of the captured step (the x frame's CFL then
# runs after the last sweep instead of inside it), mixed with fast lengths
@pytest.mark.parametrize("shape,avg", [((67, 20, 9), False), ((16, 131, 12), False), ((12, 10, 263), True),
                                       ((600, 4, 3), False), ((6, 1000, 3), False), ((4, 3, 1024), False),
                                       ((97, 67, 131), False)])
def test_ctx_steps_any_pencil_length(M, oracle, shape, avg):
    print("ragged ctx", shape, avg, flush=True)
    u0, b0 = random_state(*shape, seed=900 + sum(shape), amp=0.4)
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


def test_ctx_fp64_any_pencil_length(M, oracle):
    shape = (67, 10, 97)
    u0, b0 = random_state(*shape, seed=950, amp=0.4)
    u0, b0 = u0.astype(np.float64), b0.astype(np.float64)
    sim = M.Simulation(*shape, precision=8)
    sim.set_state(u0, b0)
    sim.step(2)
    ug, bg = sim.get_state()
    dg = sim.dts()
    sim.close()
    uo, bo = u0.copy(), b0.copy()
    do = oracle.run(uo, bo, *shape, 2)
    assert same(dg, do) and same(ug, uo) and same(bg, bo)


def test_main_drop_in_on_a_ragged_cube(M, oracle):
    n = 67
    u0, b0 = random_state(n, n, n, seed=960, amp=0.3)
    u, b = u0.copy(), b0.copy()
    dts = M.main(u, b, n, n, n, 2)
    uo, bo = u0.copy(), b0.copy()
    do = oracle.run(uo, bo, n, n, n, 2)
    assert same(dts, do) and same(u, uo) and same(b, bo)
