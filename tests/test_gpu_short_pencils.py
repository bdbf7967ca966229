"""Short pencils that share a wave: FP32 pencils of 16, 32, 48 and 96 cells
run as 4, 2, 4 and 2 pencils per wave (lane groups of 16, 32, 16 and 32
lanes, K = 1 or 3 cells per lane), with the shifts wrapping inside each
group and the pencil max of mhdflux (find_max, Heterogeneous_MHD/C/
mhdflux.c:47-57) reduced over the group only.

The shapes put those lengths in every frame, beside lengths that keep one
pencil per wave, with tiles whose last workgroup holds fewer pencils than
its lane groups (groups without a pencil), and through the fused CFL and
monitors of the last sweep.  Steps of the device-resident context, bit-exact
against the oracle: the state, every dt and the monitor history.
"""
import numpy as np
import pytest

from conftest import bitwise, same
from golden.states import random_state

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def M():
    import mhdgpu
    mhdgpu.mem_info()
    return mhdgpu


SHAPES = [
    # x 32 (2 per wave), y 20 (one), z 48 (4 per wave)
    ((32, 20, 48), False),
    # x 96 (2 per wave, K = 3), y 18, z 36 (one per wave)
    ((96, 18, 36), False),
    # x 48 (4 per wave) over 96 pencils, y 96 (2 per wave) over 48, z 30
    ((48, 96, 30), True),
    # x 16 (4 per wave) over 42 pencils: the last x tile holds 2 of 4;
    # z 32 (2 per wave) over 16 pencils
    ((16, 42, 32), False),
    # y 32 over 36 pencils (the last y/z tile half full), z 96 over 36
    ((36, 32, 96), True),
    # every frame 2 per wave at K = 3
    ((96, 96, 96), False),
]


@pytest.mark.parametrize("shape,avg", SHAPES)
def test_short_pencils_steps_match_oracle(M, oracle, shape, avg):
    print("short pencils", shape, avg, flush=True)
    nsteps = 3
    u0, b0 = random_state(*shape, seed=1300 + sum(shape), amp=0.4)
    sim = M.Simulation(*shape, avg_b=avg)
    sim.set_state(u0, b0)
    sim.step(nsteps)
    ug, bg = sim.get_state()
    dg = sim.dts()
    hist = sim.monitor_history()
    sim.close()
    uo, bo = u0.copy(), b0.copy()
    mo = []
    dts = []
    for _ in range(nsteps):
        dts.append(oracle.run(uo, bo, *shape, 1, avg_b=avg)[0])
        mo.append(oracle.monitors(uo, bo, *shape))
    assert bitwise(dg, np.array(dts, np.float32))
    assert same(ug, uo) and same(bg, bo)
    for s in range(nsteps):
        assert np.allclose(hist[s, :5], mo[s][:5], rtol=1e-13, atol=1e-9), s
        assert hist[s, 5] == mo[s][5], s


@pytest.mark.parametrize("shape", [(32, 16, 48), (96, 48, 16)])
def test_short_pencils_step_level_ops_match_oracle(M, oracle, shape):
    # the step-level AoS entry points keep one pencil per wave; the same
    # state through both paths agrees with the oracle
    print("short pencils, step level", shape, flush=True)
    u0, b0 = random_state(*shape, seed=1400 + shape[0], amp=0.4)
    dt = np.float32(2e-3)
    for op in ("fluidx", "advectbyzx"):
        du, db = M.DeviceArray.from_host(u0), M.DeviceArray.from_host(b0)
        getattr(M, op)(du, db, *shape, M.DeviceArray.from_host(np.array([dt], np.float32)))
        uo, bo = u0.copy(), b0.copy()
        getattr(oracle, op)(uo, bo, *shape, dt)
        assert same(du.to_host(), uo) and same(db.to_host(), bo), op


def test_batch_short_pencil_groups_equal_alone_and_oracle(M, oracle):
    # nine 48^3 boxes together have the pencils to fill the GPU at 4 per wave
    # (K = 3), which a lone 48^3 box does not: the batch then runs kernels the
    # lone box never uses, and every box must still equal its lone run; the
    # 32^3 and 16^3 boxes share waves at K = 1, and the 96x48x16 box has a
    # different geometry per frame
    from conftest import bitwise as bw
    dims = [(48, 48, 48)] * 9 + [(32, 32, 32)] * 2 + [(96, 48, 16), (16, 16, 16)]
    states = [random_state(*d, seed=1500 + i, amp=0.35) for i, d in enumerate(dims)]
    bt = M.Batch(dims)
    for i, (u, b) in enumerate(states):
        bt.set_state(i, u, b)
    nsteps = 3
    bt.step(nsteps)
    bt.sync()
    for i, d in enumerate(dims):
        print("batch box", i, d, flush=True)
        ug, bg = bt.get_state(i)
        sim = M.Simulation(*d)
        sim.set_state(*states[i])
        sim.step(nsteps)
        ua, ba = sim.get_state()
        da = sim.dts()
        sim.close()
        assert bw(ug, ua) and bw(bg, ba), i
        assert bw(bt.dts(i), da), i
        if i in (0, 9, 11):
            uo, bo = (a.copy() for a in states[i])
            do = oracle.run(uo, bo, *d, nsteps)
            assert bw(da, do) and same(ug, uo) and same(bg, bo), i
    bt.close()
