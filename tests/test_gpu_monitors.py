"""The FP64 conservation monitors recorded inside the captured step (total
mass, total energy, momentum x, y, z in FP64 sums; max |div B| on the left
faces), through the C ABI:

* after N steps the history holds N rows, row s the monitors of the state
  after step s+1: equal to the oracle's monitors of the oracle's state after
  the same steps (sums to FP64 rounding, the divergence exactly: a max of
  exactly formed values), and to the on-demand mhd_ctx_monitors;
* the history is bit-identical from run to run (fixed summation order);
* the fused form (inside the last sweep) and the standalone form (a ragged x
  pencil length, whose sweeps have no fused epilogue) agree the same way.
The oracle's monitors are pinned by closed forms (tests/test_oracle_pins.py).
"""
import numpy as np
import pytest

from conftest import bitwise
from golden.states import random_state

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def M():
    import mhdgpu
    mhdgpu.mem_info()
    return mhdgpu


@pytest.mark.parametrize("shape,avg", [((32, 24, 16), False), ((48, 32, 20), True), ((67, 12, 10), False)])
def test_monitor_history_matches_oracle_per_step(M, oracle, shape, avg):
    print("monitor history", shape, avg, flush=True)
    nsteps = 4
    u0, b0 = random_state(*shape, seed=1100 + shape[0], amp=0.4)
    sim = M.Simulation(*shape, avg_b=avg)
    sim.set_state(u0, b0)
    ondemand = []
    for _ in range(nsteps):
        sim.step(1)
        ondemand.append(sim.monitors())
    hist = sim.monitor_history()
    sim.close()
    assert hist.shape == (nsteps, 6)
    uo, bo = u0.copy(), b0.copy()
    for s in range(nsteps):
        oracle.run(uo, bo, *shape, 1, avg_b=avg)
        mo = oracle.monitors(uo, bo, *shape)
        assert np.allclose(hist[s, :5], mo[:5], rtol=1e-13, atol=1e-9), s
        assert hist[s, 5] == mo[5], s
        assert np.allclose(hist[s], ondemand[s], rtol=1e-13, atol=1e-9), s
        assert hist[s, 5] == ondemand[s][5]


def test_monitor_history_is_deterministic_and_conserves(M):
    shape = (64, 40, 24)
    u0, b0 = random_state(*shape, seed=1200, amp=0.3)
    runs = []
    for _ in range(2):
        sim = M.Simulation(*shape)
        sim.set_state(u0, b0)
        m0 = sim.monitors()
        sim.step(6)
        runs.append(sim.monitor_history())
        sim.reset_clock()
        assert len(sim.monitor_history()) == 0  # a reset starts a new history
        sim.close()
    assert bitwise(runs[0], runs[1])
    h = runs[0]
    assert h.shape == (6, 6)
    assert np.all(np.abs(h[:, 0] - m0[0]) <= 1e-6 * m0[0])  # mass
    assert np.all(np.abs(h[:, 1] - m0[1]) <= 1e-5 * m0[1])  # energy
    assert np.all(np.abs(h[:, 5] - m0[5]) < 1e-5)  # CT keeps div B
