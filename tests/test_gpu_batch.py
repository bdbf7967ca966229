"""Batched form: many uneven boxes advanced together must give, box by box,
exactly what each box gives alone (same dt sequence, same state), under
every schedule, and after the autotuner's trial steps have been undone."""
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


def alone(M, dims, u0, b0, steps):
    sim = M.Simulation(*dims)
    sim.set_state(u0, b0)
    sim.step(steps)
    u, b = sim.get_state()
    d = sim.dts()
    sim.close()
    return u, b, d


DIMS = [(32, 32, 32), (48, 48, 48), (24, 16, 32), (64, 64, 64), (40, 36, 28), (16, 16, 16), (96, 32, 48)]


@pytest.mark.parametrize("policy,bpc", [(0, 0), (1, 2), (2, 0), (3, 0)])
def test_batch_equals_each_box_alone(M, policy, bpc):
    states = [random_state(*d, seed=900 + i, amp=0.35) for i, d in enumerate(DIMS)]
    bt = M.Batch(DIMS)
    bt.set_schedule(bpc, policy)
    for i, (u, b) in enumerate(states):
        bt.set_state(i, u, b)
    bt.step(4)
    for i, d in enumerate(DIMS):
        print("box", i, d, flush=True)
        ug, bg = bt.get_state(i)
        ua, ba, da = alone(M, d, *states[i], 4)
        assert bitwise(bt.dts(i), da)
        assert bitwise(ug, ua) and bitwise(bg, ba)
    bt.close()


def test_batch_matches_oracle_and_autotune_restores(M, oracle):
    from mhdgpu.ics import batch_box
    dims = [(32, 32, 32), (48, 48, 48), (64, 64, 64)]
    states = [batch_box(d[0], seed=50 + i) for i, d in enumerate(dims)]
    bt = M.Batch(dims)
    for i, (u, b) in enumerate(states):
        bt.set_state(i, u, b)
    tuned = bt.autotune(trial_steps=2)
    assert tuned["policy"] in (0, 1, 2, 3) and tuned["ms_per_step"] > 0
    bt.step(3)
    for i, d in enumerate(dims):
        uo, bo = states[i][0].copy(), states[i][1].copy()
        do = oracle.run(uo, bo, *d, 3)
        ug, bg = bt.get_state(i)
        assert bitwise(bt.dts(i), do)
        assert same(ug, uo) and same(bg, bo)
    bt.close()


def test_batch_streamed_and_tiled_boxes_equal_alone(M):
    # the y/z advectbyzx streams for a box with 2048 waves of its own (192^3)
    # and for the boxes of a group that has them together: 144^3 beside it
    # (K = 3), sixteen 64^3 boxes and 40x36x28 (K = 1); 96^3, alone in its
    # group (K = 2), runs tiles.  Alone, every box but 192^3 runs tiles.
    dims = [(192, 192, 192), (144, 144, 144), (96, 96, 96)] + [(64, 64, 64)] * 16 + [(40, 36, 28)]
    states = [random_state(*d, seed=950 + i, amp=0.3) for i, d in enumerate(dims)]
    bt = M.Batch(dims)
    bt.set_schedule(0, 3)
    for i, (u, b) in enumerate(states):
        bt.set_state(i, u, b)
    bt.step(3)
    for i, d in enumerate(dims):
        print("box", i, d, flush=True)
        ug, bg = bt.get_state(i)
        ua, ba, da = alone(M, d, *states[i], 3)
        assert bitwise(bt.dts(i), da)
        assert bitwise(ug, ua) and bitwise(bg, ba)
    bt.close()


def test_batch_fused_stream_mixed_widths_equal_alone(M):
    # one group (cells per lane 8, 2, 2) whose boxes are 8 and 7 pencil tiles
    # wide: the fused y/z stream runs blocks of 8 waves, and for the narrower
    # box the eighth wave ends at once; then a state is set again between
    # steps (the CFL accumulator of that box must be formed anew)
    dims = [(512, 96, 96), (448, 96, 96)]
    states = [random_state(*d, seed=970 + i, amp=0.3) for i, d in enumerate(dims)]
    bt = M.Batch(dims)
    bt.set_schedule(0, 3)
    for i, (u, b) in enumerate(states):
        bt.set_state(i, u, b)
    bt.step(2)
    u1, b1 = bt.get_state(1)
    bt.set_state(1, u1, b1)
    bt.step(1)
    for i, d in enumerate(dims):
        print("box", i, d, flush=True)
        ug, bg = bt.get_state(i)
        ua, ba, da = alone(M, d, *states[i], 3)
        assert bitwise(bt.dts(i), da)
        assert bitwise(ug, ua) and bitwise(bg, ba)
    bt.close()
