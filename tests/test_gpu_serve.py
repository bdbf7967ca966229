"""Served form: requests that arrive while others run are admitted, batched
with them, and each result is exactly what the request gives alone; the HBM
budget queues what does not fit and refuses what never fits."""
import time

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


REQS = [((32, 32, 32), 7), ((48, 48, 48), 3), ((24, 16, 32), 5), ((64, 64, 64), 10), ((40, 36, 28), 2),
        ((16, 16, 16), 6)]


def alone(M, dims, steps, u0, b0):
    sim = M.Simulation(*dims)
    sim.set_state(u0, b0)
    sim.step(steps)
    out = sim.get_state()
    sim.close()
    return out


def run_server(M, mode, budget=0, gap_s=0.002):
    states = [random_state(*d, seed=1200 + i, amp=0.3) for i, (d, _) in enumerate(REQS)]
    sv = M.Server(mode=mode, hbm_budget=budget)
    outs, ids = {}, []
    for i, (d, n) in enumerate(REQS):
        rid, out = sv.submit(d, n, *states[i])
        outs[rid] = (i, out)
        ids.append(rid)
        time.sleep(gap_s)  # arrivals spread over the run: later ones join running ones
    done = {}
    t0 = time.time()
    while len(done) < len(REQS) and time.time() - t0 < 60:
        for r in sv.poll():
            done[r["id"]] = r
        time.sleep(0.002)
    stats = sv.stats()
    sv.close()
    assert len(done) == len(REQS)
    for rid, (i, (ou, ob)) in outs.items():
        assert done[rid]["rc"] == 0
        print("request", i, REQS[i], flush=True)
        ua, ba = alone(M, REQS[i][0], REQS[i][1], *states[i])
        assert bitwise(ou, ua) and bitwise(ob, ba)
    return done, stats


def test_server_results_equal_alone(M):
    done, stats = run_server(M, mode=0)
    assert stats["zone_updates"] == sum(np.prod(d) * n for d, n in REQS)


def test_server_budget_queues_and_refuses(M):
    # room for the 64^3 request plus a little: the others wait their turn
    budget = 64 * 64 ** 3 + 16 * 32 ** 3 + (2 << 20)
    done, stats = run_server(M, mode=0, budget=budget, gap_s=0.0)
    assert max(r["queue_s"] for r in done.values()) > 0
    sv = M.Server(mode=0, hbm_budget=budget)
    u, b = random_state(96, 96, 96, seed=1, amp=0.1)
    with pytest.raises(M.MHDError) as e:
        sv.submit((96, 96, 96), 1, u, b)
    assert e.value.code == -7
    sv.close()


def test_fifo_baseline_results_equal_alone(M):
    run_server(M, mode=1)


def test_served_poisson_mix_matches_oracle(M, oracle):
    """The served workload, reduced: seeded Poisson arrivals of requests with
    N drawn from {32, 48, 64, 96}, random-phase initial states with their own
    seeds and 10-40 steps each, continuous batching; every result equals the
    oracle's run of that request (the sign of exact zeros aside)."""
    from conftest import same
    from mhdgpu.ics import random_phase
    rng = np.random.default_rng(55)
    nreq = 12
    sizes = rng.choice(np.array([32, 48, 64, 96]), nreq).tolist()
    steps = rng.integers(10, 41, nreq).tolist()
    gaps = rng.exponential(1.0 / 200.0, nreq)  # mean 200 requests/s
    states = [random_phase(n, seed=5000 + i) for i, n in enumerate(sizes)]
    sv = M.Server(mode=0)
    outs = {}
    t0 = time.perf_counter()
    arrive = np.cumsum(gaps)
    for i, n in enumerate(sizes):
        while time.perf_counter() - t0 < arrive[i]:
            time.sleep(0.0005)
        rid, out = sv.submit((n, n, n), steps[i], *states[i])
        outs[rid] = (i, out)
    done = {}
    while len(done) < nreq and time.perf_counter() - t0 < 120:
        for r in sv.poll():
            done[r["id"]] = r
        time.sleep(0.002)
    stats = sv.stats()
    sv.close()
    assert len(done) == nreq
    assert stats["zone_updates"] == sum(n ** 3 * k for n, k in zip(sizes, steps))
    for rid, (i, (ou, ob)) in outs.items():
        assert done[rid]["rc"] == 0
        n = sizes[i]
        print("served request", i, n, steps[i], flush=True)
        uo, bo = (a.copy() for a in states[i])
        oracle.run(uo, bo, n, n, n, steps[i])
        assert same(ou, uo) and same(ob, bo), i
