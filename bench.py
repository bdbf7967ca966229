"""Benchmark of the relaxing-TVD MHD timestep on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]

Workload (the headline configuration): a 256^3 periodic FP32 box, the
Alfven background (rho0 = 1, p0 = 0.6, B0 = x) plus a seeded random-phase
velocity field (P(k) ~ k^-2, k <= 8, rms Mach 0.1) and a curl-A magnetic
perturbation (rms 0.1 B0).  One step = one loop iteration of the reference
(1 CFL reduction + dt + 12 sweeps, advancing t by 2 dt).  The state is
resident in HBM before the timed region; the K timed steps are replays of
the captured step graph, bracketed by a device synchronisation on both
sides.  Prints ONE JSON line.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "mhd-gpu_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

HBM_PEAK_GBS = 8000.0
FLUIDX_BYTES_PER_CELL = 52   # u read 20 + b read 12 + u write 20
ADVECT_BYTES_PER_CELL = 32   # u1,u2 read 8 + b read 12 + b write 12
CFL_BYTES_PER_CELL = 32      # u, b read
STEP_BYTES_PER_CELL = 6 * FLUIDX_BYTES_PER_CELL + 6 * ADVECT_BYTES_PER_CELL + CFL_BYTES_PER_CELL  # 536
SWEEPS = ["fluidx_x", "advect_x", "fluidx_y", "advect_y", "fluidx_z", "advect_z", "advect_z", "fluidx_z",
          "advect_y", "fluidx_y", "advect_x", "fluidx_x"]


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def profile_value(fname, key):
    """A `key: value` line of a committed rocprofv3 summary under profiles/, if any."""
    path = os.path.join(ROOT, "profiles", fname)
    if not os.path.exists(path):
        return None
    for line in open(path):
        line = line.strip()
        if line.startswith(key + ":"):
            return float(line.split(":")[1].split()[0])
    return None


def pmc_traffic(kernel="fluidx", n=256):
    """HBM bytes per launch from the committed rocprofv3 --pmc summary, if any."""
    return profile_value(f"{kernel}_{n}cube_pmc.txt", "traffic_bytes_per_launch")


def valu_roofline(fl_ms, n):
    """fluidx is limited by VALU issue, not HBM: its VALU operations per cell
    (rocprofv3 SQ counters, committed summary) at 4 cycles per wave64
    operation on 256 CUs x 4 SIMDs at 2.4 GHz give a lower bound on the time
    of a launch (100% VALU issue, transcendentals counted at full rate)."""
    vpc = profile_value(f"sweeps_{n}cube_sq.txt", "valu_per_cell_fluidx")
    if vpc is None:
        return None
    cyc = vpc / 64 * 4
    floor_ms = n ** 3 * cyc / (256 * 4) / 2.4e9 * 1e3
    return {"kernel": "fluidx", "bound": "valu", "valu_ops_per_cell": vpc, "simd_cycles_per_cell": cyc,
            "floor_ms": floor_ms, "avg_launch_ms": fl_ms, "frac": floor_ms / fl_ms}


def kernels_hbm(sweep_ms, cells):
    """Every sweep kernel of the step against the HBM peak: algorithmic bytes
    per launch (each field read once, each updated field written once) over
    its HIP-event launch time (mhd_ctx_profile_step)."""
    fam = {"fluidx_x": ([0], FLUIDX_BYTES_PER_CELL), "fluidx_x_with_cfl_and_monitors": ([11], FLUIDX_BYTES_PER_CELL),
           "fluidx_yz": ([2, 4, 7, 9], FLUIDX_BYTES_PER_CELL), "advect_x": ([1, 10], ADVECT_BYTES_PER_CELL),
           "advect_yz": ([3, 5, 6, 8], ADVECT_BYTES_PER_CELL)}
    out = {}
    for k, (idx, bpc) in fam.items():
        ms = float(np.mean([sweep_ms[i] for i in idx]))
        gbs = bpc * cells / (ms * 1e-3) / 1e9
        out[k] = {"bytes_per_cell": bpc, "avg_launch_ms": ms, "achieved_GBs": gbs, "frac": gbs / HBM_PEAK_GBS}
    return out


def host_info():
    """The host the CPU baselines ran on: logical CPUs this process may use,
    the lscpu model and socket / core / thread counts, OMP_NUM_THREADS."""
    import subprocess
    info = {"nproc": len(os.sched_getaffinity(0)), "os_cpu_count": os.cpu_count(),
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        keys = {"Model name": "model", "Socket(s)": "sockets", "Core(s) per socket": "cores_per_socket",
                "Thread(s) per core": "threads_per_core", "NUMA node(s)": "numa_nodes"}
        for line in out.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in keys:
                v = v.strip()
                info[keys[k.strip()]] = int(v) if v.isdigit() else v
        if "sockets" in info and "cores_per_socket" in info:
            info["physical_cores"] = info["sockets"] * info["cores_per_socket"]
    except (OSError, subprocess.SubprocessError):
        pass
    return info


def cpu_baselines(u0, b0, n):
    """The CPU path beside the GPU one (reported baselines, not the target):
    * the reference's own C path (oracle/_ref, compiled from its sources;
      serial, as the reference runs) on 1 step of the n^3 workload (about
      14 s at 256^3) and on the 32^3 linear Alfven workload in full (200
      steps); where oracle/_ref is absent, the oracle restatement instead;
    * the oracle restatement (the reference C path's evaluation order) with
      its pencil loops spread over OpenMP threads, as the reference's Fortran
      OpenMP driver spreads them (OpenMP/mhd.f90), on 3 steps of n^3."""
    import oracle_lib
    from mhdgpu.ics import alfven_linear
    if oracle_lib.reference_available():
        serial, kind, cores = oracle_lib.Reference(), "reference", 1
    else:
        serial, kind, cores = oracle_lib.Oracle(), "port", oracle_lib.Oracle().threads()
    out = {}
    u, b = u0.copy(), b0.copy()
    t0 = time.perf_counter()
    serial.run(u, b, n, n, n, 1)
    dt = time.perf_counter() - t0
    out["cpu_baseline"] = {"value": n ** 3 / dt, "unit": "zone-updates/s", "cores": cores, "kind": kind,
                           "sample": f"1 step of the {n}^3 workload (same initial state), {dt:.1f} s wall"}
    u, b = alfven_linear(32)
    t0 = time.perf_counter()
    serial.run(u, b, 32, 32, 32, 200)
    dt = time.perf_counter() - t0
    out["cpu_baseline_config1"] = {"value": 32 ** 3 * 200 / dt, "unit": "zone-updates/s", "cores": cores,
                                   "kind": kind, "sample": f"32^3 linear Alfven wave, 200 steps in full, {dt:.1f} s wall"}
    orc = oracle_lib.Oracle()
    u, b = u0.copy(), b0.copy()
    t0 = time.perf_counter()
    orc.run(u, b, n, n, n, 3)
    dt = time.perf_counter() - t0
    out["cpu_baseline_omp"] = {"value": 3 * n ** 3 / dt, "unit": "zone-updates/s", "cores": orc.threads(),
                               "kind": "port", "sample": f"3 steps of the {n}^3 workload (same initial state), "
                                                         f"{dt:.2f} s wall, OpenMP over pencils"}
    out["cpu_host"] = host_info()
    return out


def bench_batched(mhdgpu, steps=100, nbox=64, seed=3):
    """The batched workload: nbox independent boxes, N uniform over
    {32,48,64,96,128} (seeded), each the linear Alfven wave with its own seeded
    phase and amplitude, `steps` steps each.  Compared with the same boxes run
    one after the other through the plain path (one context per box)."""
    from mhdgpu.ics import batch_box, batch_sizes
    sizes = batch_sizes(nbox, seed)
    states = [batch_box(n, seed=1000 + i) for i, n in enumerate(sizes)]
    cells = sum(n ** 3 for n in sizes)
    log(f"bench: batched, {nbox} boxes, {cells / 1e6:.1f} M cells")
    # sequential plain runs (graph captured before each timed region)
    seq_ms, ref_states = 0.0, {}
    for i, n in enumerate(sizes):
        sim = mhdgpu.Simulation(n, n, n)
        sim.set_state(*states[i])
        seq_ms += sim.step_timed(steps)
        if i < 3:
            ref_states[i] = sim.get_state()
        sim.close()
    # batched: autotune on the batch itself (state restored), then time
    bt = mhdgpu.Batch([(n, n, n) for n in sizes])
    for i in range(nbox):
        bt.set_state(i, *states[i])
    tuned = bt.autotune(trial_steps=2)
    t0 = time.perf_counter()
    bat_ms = bt.step_timed(steps)
    mhdgpu.check(mhdgpu.lib().mhd_device_sync())
    wall = time.perf_counter() - t0
    same = all(np.array_equal(bt.get_state(i)[0], ref_states[i][0]) and
               np.array_equal(bt.get_state(i)[1], ref_states[i][1]) for i in ref_states)
    bt.close()
    zu = cells * steps
    return {"boxes": nbox, "sizes_hist": {str(k): sizes.count(k) for k in sorted(set(sizes))},
            "cells": cells, "steps": steps,
            "batched_zone_updates_per_s": zu / (bat_ms * 1e-3), "batched_ms": bat_ms, "batched_wall_ms": wall * 1e3,
            "sequential_zone_updates_per_s": zu / (seq_ms * 1e-3), "sequential_ms": seq_ms,
            "speedup_vs_sequential": seq_ms / bat_ms, "schedule": tuned,
            "boxes_identical_to_plain": bool(same)}


def served_workload(n_req=64, seed=5, rate=20.0):
    """Poisson arrivals (mean `rate` req/s, seeded) of requests with N uniform
    over {32,48,64,96,128,192} and steps uniform in [50,500]; each initial
    state is the random-phase state of its size, cyclically shifted by
    seeded offsets (a distinct periodic state per request)."""
    from mhdgpu.ics import random_phase
    rng = np.random.default_rng(seed)
    sizes = rng.choice(np.array([32, 48, 64, 96, 128, 192]), n_req)
    steps = rng.integers(50, 501, n_req)
    gaps = rng.exponential(1.0 / rate, n_req)
    arrivals = np.cumsum(gaps) - gaps[0]
    base = {int(n): random_phase(int(n), seed=int(n)) for n in sorted(set(sizes.tolist()))}
    states = []
    for i, n in enumerate(sizes.tolist()):
        sh = rng.integers(0, n, 3)
        u, b = base[n]
        u = np.roll(u.reshape(n, n, n, 5), tuple(sh), axis=(0, 1, 2)).ravel()
        b = np.roll(b.reshape(n, n, n, 3), tuple(sh), axis=(0, 1, 2)).ravel()
        states.append((u, b))
    return sizes.tolist(), steps.tolist(), arrivals.tolist(), states


def run_served(mhdgpu, mode, sizes, steps, arrivals, states):
    sv = mhdgpu.Server(mode=mode)
    t0 = time.perf_counter()
    done = {}
    for i in range(len(sizes)):
        while time.perf_counter() - t0 < arrivals[i]:
            for r in sv.poll():
                done[r["id"]] = r
            time.sleep(0.0005)
        n = sizes[i]
        sv.submit((n, n, n), steps[i], *states[i])
    while len(done) < len(sizes):
        for r in sv.poll():
            done[r["id"]] = r
        time.sleep(0.001)
    st = sv.stats()
    sv.close()
    lat = np.array([r["latency_s"] for r in done.values()])
    ok = all(r["rc"] == 0 for r in done.values())
    return {"p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "mean_s": float(lat.mean()), "zone_updates": st["zone_updates"], "span_s": st["span_s"],
            "zone_updates_per_s": st["zone_updates"] / st["span_s"], "gpu_busy_s": st["gpu_busy_s"],
            "rebuilds": int(st["rebuilds"]), "all_ok": ok,
            "service_thread_s": {k[2:-2]: st[k] for k in ("t_upload_s", "t_rebuild_s", "t_step_s", "t_download_s")}}


def bench_served(mhdgpu, n_req=64):
    sizes, steps, arrivals, states = served_workload(n_req)
    log(f"bench: served, {n_req} requests over {arrivals[-1]:.2f} s")
    batched = run_served(mhdgpu, 0, sizes, steps, arrivals, states)
    log("bench: served FIFO baseline")
    fifo = run_served(mhdgpu, 1, sizes, steps, arrivals, states)
    return {"requests": n_req, "arrival_rate_per_s": 20.0, "arrival_span_s": arrivals[-1],
            "offered_zone_updates": float(sum(n ** 3 * k for n, k in zip(sizes, steps))),
            "continuous_batching": batched, "fifo_one_at_a_time": fifo,
            "throughput_ratio": batched["zone_updates_per_s"] / fifo["zone_updates_per_s"]}


VARS = ["rho", "mx", "my", "mz", "E", "bx", "by", "bz"]


def precision_report(mhdgpu, n, steps, ic, seed, state=None):
    """FP32 against FP64 on the device from the same initial state: relative
    L2 error per variable after `steps` steps, FP64 drift of the total mass
    and energy, max |div B| (left faces), and the dt of both runs.  state:
    the initial (u, b) when the caller has generated it already."""
    from mhdgpu.ics import random_phase, turbulence
    if state is not None:
        u0, b0 = state
    elif ic == "turbulence":
        log(f"precision: generating {ic} {n}^3 (seed {seed})")
        u0, b0 = turbulence(n, seed=seed, progress=log)
    else:
        log(f"precision: generating {ic} {n}^3 (seed {seed})")
        u0, b0 = random_phase(n, seed=seed)
    runs = {}
    for prec in (8, 4):
        name = "fp64" if prec == 8 else "fp32"
        sim = mhdgpu.Simulation(n, n, n, precision=prec)
        sim.set_state(u0, b0)
        m0 = sim.monitors()
        log(f"precision: {name} {steps} steps")
        t0 = time.perf_counter()
        done = 0
        while done < steps:
            k = min(50, steps - done)
            sim.step(k)
            sim.sync()
            done += k
            log(f"precision: {name} {done}/{steps}")
        wall = time.perf_counter() - t0
        m1 = sim.monitors()
        d = sim.dts()
        u, b = sim.get_state()
        t, t64 = sim.time()
        sim.close()
        runs[name] = {"u": u, "b": b, "m0": m0, "m1": m1, "dts": d, "ms_per_step": wall * 1e3 / steps,
                      "t": t64}
    r64, r32 = runs["fp64"], runs["fp32"]
    rel = {}
    for c in range(5):
        a, e = r32["u"][c::5].astype(np.float64), r64["u"][c::5]
        rel[VARS[c]] = float(np.linalg.norm(a - e) / np.linalg.norm(e))
    for c in range(3):
        a, e = r32["b"][c::3].astype(np.float64), r64["b"][c::3]
        rel[VARS[5 + c]] = float(np.linalg.norm(a - e) / np.linalg.norm(e))
    allv = (sum(float(np.sum((r32["u"][c::5] - r64["u"][c::5]) ** 2)) for c in range(5)) +
            sum(float(np.sum((r32["b"][c::3] - r64["b"][c::3]) ** 2)) for c in range(3)))
    alln = float(np.sum(r64["u"] ** 2) + np.sum(r64["b"] ** 2))
    out = {"mode": "precision", "ic": ic, "grid": [n, n, n], "steps": steps, "seed": seed,
           "rel_l2_fp32_vs_fp64": rel, "rel_l2_all": (allv / alln) ** 0.5}
    for name, r in runs.items():
        out[name] = {"mass_drift": float((r["m1"][0] - r["m0"][0]) / r["m0"][0]),
                     "energy_drift": float((r["m1"][1] - r["m0"][1]) / r["m0"][1]),
                     "max_divB_start": float(r["m0"][5]), "max_divB_end": float(r["m1"][5]),
                     "t_end": r["t"], "dt_first": float(r["dts"][0]), "dt_last": float(r["dts"][-1]),
                     "finite": bool(np.isfinite(r["u"]).all() and np.isfinite(r["b"]).all()),
                     "ms_per_step": r["ms_per_step"]}
    return out


def steplevel_report(mhdgpu, n, steps, warmup, state=None):
    """The reference-shaped path: the loop of HIP/cu/cuda_main.cu:89-121 driven
    through the step-level entry points on AoS device arrays (mhd_cfl, then
    dt = 0.9*cfl clamped as Heterogeneous_MHD/C/mhd.c:63-65 on the host, then
    fluidx / advectbyzx with the permuted dims and the four transposes, which
    write ut, bt and copy back).  Every entry point is synchronous, like the
    reference's wrappers; per-op times are host wall times around each call."""
    from mhdgpu.ics import random_phase
    u0, b0 = state if state is not None else random_phase(n, seed=2)
    du, db = mhdgpu.DeviceArray.from_host(u0), mhdgpu.DeviceArray.from_host(b0)
    dut, dbt = mhdgpu.DeviceArray(u0.nbytes), mhdgpu.DeviceArray(b0.nbytes)
    dcfl, ddt = mhdgpu.DeviceArray(4), mhdgpu.DeviceArray(4)
    acc = {}

    def op(name, fn, *a):
        t0 = time.perf_counter()
        fn(*a)
        acc[name] = acc.get(name, 0.0) + time.perf_counter() - t0

    M = mhdgpu
    X = (n, n, n)
    seq = [("fluidx", X), ("advect", X), ("t12", X), ("fluidx", X), ("advect", X), ("t13", X), ("fluidx", X),
           ("advect", X), ("advect", X), ("fluidx", X), ("t13", X), ("advect", X), ("fluidx", X), ("t12", X),
           ("advect", X), ("fluidx", X)]

    def step():
        t0 = time.perf_counter()
        c = M.cfl(du, db, n, n, n, dcfl)
        dt = np.float32(0.9 * float(np.float32(c)))
        ddt.copy_from(np.array([dt], np.float32))
        acc["cfl+dt"] = acc.get("cfl+dt", 0.0) + time.perf_counter() - t0
        for name, d in seq:
            if name == "fluidx":
                op(name, M.fluidx, du, db, *d, ddt)
            elif name == "advect":
                op(name, M.advectbyzx, du, db, *d, ddt)
            elif name == "t12":
                op(name, M.transpose12, dut, dbt, du, db, *d)
            else:
                op(name, M.transpose13, dut, dbt, du, db, *d)

    for _ in range(warmup):
        step()
    acc.clear()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    wall = time.perf_counter() - t0
    counts = {"fluidx": 6, "advect": 6, "t12": 2, "t13": 2, "cfl+dt": 1}
    # algorithmic HBM bytes per cell of each call on the AoS arrays: fluidx
    # reads u and b and writes u; advectbyzx copies b, then reads u (whole
    # 20-byte records), b and writes b; a transpose reads u, b and writes
    # both buffers; the CFL reads u and b
    bpc = {"fluidx": 52, "advect": 24 + 44, "t12": 32 + 64, "t13": 32 + 64, "cfl+dt": 32}
    per = {k: v * 1e3 / (steps * counts[k]) for k, v in acc.items()}
    return {"mode": "steplevel", "grid": [n, n, n], "steps": steps, "ms_per_step": wall * 1e3 / steps,
            "zone_updates_per_s": n ** 3 * steps / wall,
            "ms_per_call": per,
            "hbm_GBs_per_call": {k: bpc[k] * n ** 3 / (per[k] * 1e-3) / 1e9 for k in per},
            "note": "synchronous calls; host wall time per call includes the launch and the sync"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-served", action="store_true")
    ap.add_argument("--no-precision", action="store_true")
    ap.add_argument("--no-steplevel", action="store_true")
    ap.add_argument("--requests", type=int, default=64)
    ap.add_argument("--mode", choices=["plain", "copy", "precision", "steplevel", "selfcheck"], default="plain",
                    help="copy: time the plain copy kernel on 1 GiB (counter calibration); "
                         "precision: FP32 vs FP64 report on --ic at --n for --steps; "
                         "steplevel: the reference's loop through the step-level AoS entry points; "
                         "selfcheck: the exhaustive scans of the reciprocal instructions the exact "
                         "FP32 division and limiter shortcuts rely on, and of the shortened square "
                         "root of float-valued arguments")
    ap.add_argument("--ic", choices=["random_phase", "turbulence"], default="random_phase")
    ap.add_argument("--seed", type=int, default=2)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: the library runs on one GPU only; runs on several GPUs are not supported", file=sys.stderr)
        sys.exit(2)

    import mhdgpu
    from mhdgpu.ics import random_phase

    if args.mode == "copy":
        nbytes = 1 << 30
        src, dst = mhdgpu.DeviceArray(nbytes, np.uint8), mhdgpu.DeviceArray(nbytes, np.uint8)
        for _ in range(args.warmup):
            mhdgpu.check(mhdgpu.lib().mhd_dev_copy_kernel(dst.ptr, src.ptr, nbytes))
        t0 = time.perf_counter()
        for _ in range(args.steps):
            mhdgpu.check(mhdgpu.lib().mhd_dev_copy_kernel(dst.ptr, src.ptr, nbytes))
        wall = time.perf_counter() - t0
        print(json.dumps({"mode": "copy", "bytes_copied": nbytes, "launches": args.steps,
                          "GBs_read_plus_write": 2 * nbytes * args.steps / wall / 1e9}), flush=True)
        return

    if args.mode == "steplevel":
        print(json.dumps(steplevel_report(mhdgpu, args.n, args.steps, args.warmup)), flush=True)
        return

    if args.mode == "selfcheck":
        t0 = time.perf_counter()
        rcp64 = mhdgpu.selfcheck()
        bad_q, bad_rcp = mhdgpu.selfcheck_division(1)
        bad_sqrt = mhdgpu.selfcheck_sqrt()
        print(json.dumps({"mode": "selfcheck", "rcp_f64_max_rel_err": rcp64,
                          "quotient_pairs": 1 << 46, "one_correction_quotient_mismatches": bad_q,
                          "rcp_f32_equivariance_mismatches": bad_rcp,
                          "float_valued_sqrt_mismatches": bad_sqrt,
                          "wall_s": time.perf_counter() - t0}), flush=True)
        return

    if args.mode == "precision":
        print(json.dumps(precision_report(mhdgpu, args.n, args.steps, args.ic, args.seed)), flush=True)
        return

    n = args.n
    log(f"bench: generating {n}^3 initial state")
    u0, b0 = random_phase(n, seed=2)
    sim = mhdgpu.Simulation(n, n, n)
    sim.set_state(u0, b0)
    log("bench: warmup")
    sim.step(args.warmup)
    sim.sync()
    mhdgpu.check(mhdgpu.lib().mhd_device_sync())
    t0 = time.perf_counter()
    ev_ms = sim.step_timed(args.steps)
    mhdgpu.check(mhdgpu.lib().mhd_device_sync())
    wall = time.perf_counter() - t0
    ms_per_step = 1e3 * wall / args.steps
    zus = n ** 3 * args.steps / wall
    log(f"bench: {args.steps} steps in {wall * 1e3:.1f} ms (events {ev_ms:.1f} ms) -> {zus / 1e9:.3f} G zone-updates/s")

    # per-kernel times with HIP events on the context's stream
    prof = sim.profile_step(reps=5)
    sweep_ms = prof[1:]
    # the dominant kernel is the fluidx of the y and z frames (fluidx_soa_kernel,
    # four launches per step); the x-frame fluidx is reported beside it, and
    # the last sweep (fluidx x) also forms the next step's CFL and monitors
    fl_yz = [sweep_ms[i] for i, s in enumerate(SWEEPS) if s in ("fluidx_y", "fluidx_z")]
    fl = [sweep_ms[i] for i, s in enumerate(SWEEPS) if s.startswith("fluidx") and i != 11]
    ad = [sweep_ms[i] for i, s in enumerate(SWEEPS) if s.startswith("advect")]
    fl_yz_ms = float(np.mean(fl_yz))
    fl_ms = float(np.mean(fl))
    ad_ms = float(np.mean(ad))
    cells = n ** 3
    achieved = FLUIDX_BYTES_PER_CELL * cells / (fl_yz_ms * 1e-3) / 1e9
    m = sim.monitors()
    t, t64 = sim.time()
    finite = bool(np.isfinite(m).all())
    hist = sim.monitor_history()
    mass_drift = float((hist[-1, 0] - hist[0, 0]) / hist[0, 0]) if len(hist) else None
    energy_drift = float((hist[-1, 1] - hist[0, 1]) / hist[0, 1]) if len(hist) else None
    max_divb = float(hist[-1, 5]) if len(hist) else None
    sim.close()

    out = {
        "metric": "zone-updates/s @256^3 FP32 (plain)",
        "value": zus,
        "unit": "zone-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{n}^3 periodic FP32 box, Alfven background + random-phase velocity (rms Mach 0.1) "
                               f"+ curl-A field (rms 0.1 B0), plain method, {args.steps} timed steps",
                   "grid": [n, n, n], "method": "plain", "precision": "fp32"},
        "roofline": {"kernel": "fluidx_soa_kernel (fluidx, y and z frames)", "bound": "hbm", "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": pmc_traffic("fluidx", n), "bytes_per_launch": FLUIDX_BYTES_PER_CELL * cells,
                     "avg_launch_ms": fl_yz_ms,
                     "timing": "HIP events around each launch on the context's stream (mhd_ctx_profile_step, "
                               "5 steps)",
                     "trace_avg_launch_ms": profile_value(f"step_{n}cube_kernel_trace.txt", "fluidx_yz_trace_avg_ms")},
        "valu_roofline": valu_roofline(fl_ms, n),
        "monitors_fp64_per_step": {"recorded": int(len(hist)), "mass_drift_over_timed_steps": mass_drift,
                                   "energy_drift_over_timed_steps": energy_drift, "max_divB_last": max_divb},
        "step_roofline": {"bytes_per_zone_update": STEP_BYTES_PER_CELL,
                          "achieved_GBs": zus * STEP_BYTES_PER_CELL / 1e9,
                          "frac": zus * STEP_BYTES_PER_CELL / 1e9 / HBM_PEAK_GBS},
        "kernels_hbm": kernels_hbm(sweep_ms, cells),
        "kernel_ms": {"dt": float(prof[0]), "fluidx_yz_avg": fl_yz_ms, "fluidx_avg_5": fl_ms,
                      "fluidx_x": float(sweep_ms[0]), "fluidx_x_with_cfl_and_monitors": float(sweep_ms[11]),
                      "advect_avg": ad_ms,
                      "sweeps": {f"{i:02d}_{s}": float(v) for i, (s, v) in enumerate(zip(SWEEPS, sweep_ms))}},
        "monitors_fp64": {"mass": m[0], "energy": m[1], "max_divB": m[5], "t": t, "t64": t64, "finite": finite},
    }
    # the drop-in program loop on host buffers (cuda_main's contract): the
    # rate includes device allocation, the H2D copy of u and b, the steps and
    # the D2H copy back; a first 1-step call loads the code objects
    uh, bh = u0.copy(), b0.copy()
    mhdgpu.main(uh, bh, n, n, n, 1)
    uh, bh = u0.copy(), b0.copy()
    t0 = time.perf_counter()
    mhdgpu.main(uh, bh, n, n, n, args.steps)
    wall_main = time.perf_counter() - t0
    out["drop_in_main"] = {"steps": args.steps, "wall_ms": wall_main * 1e3,
                           "zone_updates_per_s": n ** 3 * args.steps / wall_main,
                           "includes": "device allocation, H2D of u and b, the steps, D2H (PCIe-inclusive)"}
    del uh, bh
    if not args.no_steplevel:
        # the reference's own call sequence through the step-level entry points
        log("bench: step-level loop, 5 steps")
        out["steplevel"] = steplevel_report(mhdgpu, n, 5, 1, state=(u0, b0))
    if not args.no_batched:
        out["batched"] = bench_batched(mhdgpu)
    if not args.no_served:
        out["served"] = bench_served(mhdgpu, args.requests)
    if not args.no_precision:
        log("bench: precision, FP32 against FP64, 100 steps")
        pr = precision_report(mhdgpu, n, 100, "random_phase", 2, state=(u0, b0))
        worst = max(pr["rel_l2_fp32_vs_fp64"], key=pr["rel_l2_fp32_vs_fp64"].get)
        out["precision_fp32_vs_fp64"] = {
            "workload": f"the same {n}^3 state, 100 steps, FP32 and FP64 on the device",
            "rel_l2_all": pr["rel_l2_all"], "rel_l2_worst": {worst: pr["rel_l2_fp32_vs_fp64"][worst]},
            "fp32": {k: pr["fp32"][k] for k in ("mass_drift", "energy_drift", "max_divB_end", "ms_per_step")},
            "fp64": {k: pr["fp64"][k] for k in ("mass_drift", "energy_drift", "max_divB_end", "ms_per_step")}}
    if not args.no_cpu:
        log("bench: CPU baselines")
        out.update(cpu_baselines(u0, b0, n))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
