#!/usr/bin/env python3
"""Benchmark of the ABC-PMC simulation hot path on one MI355X.

Metric: ABC simulations/s and particle-steps/s, batched vs plain, with p50/p99
simulation latency. One timed step = one ABC generation of the batched
workload (100,000 proposals, concentration log-uniform on 1e8..1e10 per mL
-> 120..12,000 particles, frames uniform on {500, 1000, 2000}, H 1000 x 500
uint32) run as one captured hipGraph launch (reset, persistent walk,
distance) plus the read-back of the distances. Inputs are resident in HBM
before timing starts. Beside it:
  - plain form: 10,000 proposals of 1,000 frames, one walk + one distance
    launch per simulation, spread over 4 streams like the reference's job
    pool (and on one stream);
  - serving form: Poisson arrivals of single-simulation requests from 32
    chains with their own models, p50/p99 latency from each request's
    scheduled arrival, for the host-driven launch groups and for the
    device-resident request loop;
  - roofline of the walk kernel: algorithmic lane-operations per
    particle-step x particle-steps over the walk's duration (HIP events on
    the batch stream) against the FP32 vector peak;
  - CPU baseline: the oracle's FP64 restatement of the reference's CPU
    simulate + distance on the CPU workload (100 proposals from the prior,
    500 frames, H 500 x 200) on all the job's host threads and on one.
Prints one JSON line.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import contextlib
import ctypes
import json
import math
import os
import pathlib
import sys
import time

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))

from part_conc_spt_abc_amd import _abi  # noqa: E402
from part_conc_spt_abc_amd.gpu import Context, Server  # noqa: E402
from part_conc_spt_abc_amd.model import Model, prior_record, proposal_record, stokes_einstein  # noqa: E402

# The walk is bound by VALU issue. Its roofline counts the work the algorithm
# needs per particle-step for the random stream it draws, one lane-operation
# per FP32 or 32-bit integer operation and one each for log, sqrt, sin and cos:
#   Philox4x32-10: 2 blocks per 4-step chunk (3 with Gaussian detection), 104
#     operations a block (10 rounds of 2 32x32->64 multiplies = 4, 4 XORs and
#     2 key additions, plus the 4 counter words);
#   3 unit normals a step = 1.5 Box-Muller pairs of 12 operations (2 uniform
#     conversions of 2 operations, log, x -2, sqrt, x 2 pi, sin, cos, 2
#     products);
#   27 for the step itself: move 3 axes, wrap 3, window tests and their AND,
#     MSD accumulation, track bookkeeping;
#   Gaussian detection 6 more: a uniform conversion (2), z^2 scaled (2), exp,
#     compare.
# That count x particle-steps over the walk kernel's duration is held against
# MI355X's FP32 vector peak, 157.3 TFLOP/s. The VALU lane-instructions the
# kernel issues per step (rocprofv3 SQ_INSTS_VALU x 64, profiles/) give the
# issue-slot utilisation beside it, against 78.65e12 lane-instructions/s
# (256 CU x 128 lanes/cycle x 2.4 GHz), and the profiled HBM bytes per step
# the traffic.
PMC_PROFILE = ROOT / "profiles" / "walk_kernel_pmc.json"
FP32_PEAK = 157.3e12
ISSUE_PEAK = 157.3e12 / 2
PHILOX_BLOCK_OPS = 104


def walk_ops_per_step(gauss: bool = False, words_per_normal: float = 0.0) -> float:
    """Algorithmic lane-ops per particle-step of the implemented stream (2
    Philox blocks per 4 steps, 3 with Gaussian detection). words_per_normal
    > 0 counts a stream that draws that many Philox words per unit normal
    instead (1.0: one word each, 0.75 blocks per step, 123 ops)."""
    if words_per_normal > 0.0:
        philox = (3.0 * words_per_normal + (1.0 if gauss else 0.0)) / 4.0 * PHILOX_BLOCK_OPS
    else:
        philox = (3.0 if gauss else 2.0) / 4.0 * PHILOX_BLOCK_OPS
    normals = 1.5 * (2 * 2 + 8)
    step = 27.0
    detect = 6.0 if gauss else 0.0
    return philox + normals + step + detect

D175 = stokes_einstein(175.0)
D500 = stokes_einstein(500.0)


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def headline_model(z_boundary: str = "periodic") -> Model:
    zb = _abi.PCS_Z_REFLECT if z_boundary == "reflect" else _abi.PCS_Z_PERIODIC
    return Model(Lx=200.0, Ly=200.0, Lz=30.0, ax=160.0, ay=120.0, deltat=0.05, kmin=3, n_k=1000, n_de=500,
                 ub_de=20.0, n_species=2, z_boundary=zb)


def workload(model, n, rng, frames_choices):
    props, frames, parts = [], [], []
    lam_f = model.volume_factor
    for i in range(n):
        ct = 10.0 ** rng.uniform(8.0, 10.0)
        props.append(proposal_record(model, [D175, D500], [ct / 2, ct / 2], [2.0], i, i, 1))
        frames.append(int(rng.choice(frames_choices)))
        parts.append(int(rng.poisson(ct * lam_f)))
    return props, np.array(frames, np.uint32), np.array(parts, np.uint32)


def particle_steps(frames, parts):
    return float(np.sum(parts.astype(np.float64) * (frames.astype(np.float64) - 1.0)))


def observed_hist(ctx, model):
    """Data histogram at the truth (c = 1e9/mL, 50/50), seed 0, 1000 frames."""
    rng = np.random.default_rng(0)
    n = int(rng.poisson(1e9 * model.volume_factor))
    p = proposal_record(model, [D175, D500], [5e8, 5e8], [2.0], 0xFFFFFFF0, 0, 1)
    b = ctx.batch()
    b.load([p], np.array([1000], np.uint32), np.array([n], np.uint32))
    b.run(0, _abi.PCS_MODE_BATCHED)
    _, _, H = b.read(hist=True)
    b.close()
    return H[0]


def run_timed(batch, steps, warmup, mode, seed0):
    """Wall time per step (run + read of the distances) and the device times
    of each step from HIP events on the batch stream (whole run, walk
    kernel; the walk time only in PCS_MODE_BATCHED_DIRECT)."""
    for i in range(warmup):
        batch.run(seed0 + 1000 + i, mode)
        batch.read()
    walk_ms, total_ms, walls = [], [], []
    for i in range(steps):
        t0 = time.perf_counter()
        batch.run(seed0 + i, mode)
        batch.read()
        walls.append(time.perf_counter() - t0)
        tot, walk, _ = batch.timing()
        total_ms.append(tot)
        walk_ms.append(walk)
    return walls, total_ms, walk_ms


@contextlib.contextmanager
def precise_timers():
    """Timer slack of the calling thread at 1 ns (Linux prctl PR_SET_TIMERSLACK)
    while the load generator runs: its timed waits for the next arrival then
    end on time instead of up to 50 us late (the default slack), so requests
    are submitted when they are due. Restored to the thread's default after."""
    libc = ctypes.CDLL(None, use_errno=True)
    set_ok = libc.prctl(29, ctypes.c_ulong(1), 0, 0, 0) == 0
    try:
        yield set_ok
    finally:
        if set_ok:
            libc.prctl(29, ctypes.c_ulong(0), 0, 0, 0)


def serving_latency(n_requests, rate, seed=7, device_loop=False, one_wave_tiles=False):
    """Poisson arrivals at `rate` req/s from 32 chains, each with its own box
    (lateral 100-400 um, depth 10-50 um) and 2-4 species (100-600 nm);
    host-driven launch groups or the device-resident loop (one_wave_tiles:
    small requests' tiles walked by one wave, not a group, each). A request's
    latency runs from its scheduled arrival (t0 + the sum of the gaps before
    it, on the server's clock) to the moment its result is seen, so time it
    waits behind a busy submit loop or a full queue counts."""
    rng = np.random.default_rng(seed)
    srv = Server(device=0, arena_bytes=4 << 30, max_inflight=4096, device_loop=device_loop,
                 one_wave_tiles=one_wave_tiles)
    chains = []
    for c in range(32):
        L = float(rng.uniform(100, 400))
        Lz = float(rng.uniform(10, 50))
        C = int(rng.integers(2, 5))
        m = Model(Lx=L, Ly=L, Lz=Lz, ax=0.8 * L, ay=0.6 * L, deltat=0.05, kmin=3, n_k=1000, n_de=500, ub_de=20.0,
                  n_species=C)
        sid = srv.add_system(m)
        D = sorted(stokes_einstein(float(d)) for d in rng.uniform(100, 600, C))
        chains.append((m, sid, D))
    reqs = []
    for i in range(n_requests):
        m, sid, D = chains[i % 32]
        ct = 10.0 ** rng.uniform(8.0, 10.0)
        w = rng.dirichlet(np.ones(m.n_species))
        p = proposal_record(m, D, list(w * ct), [2.0], i, 0, 1)
        n = int(rng.poisson(ct * m.volume_factor))
        reqs.append((sid, p, np.array([1000], np.uint32), np.array([min(n, 60000)], np.uint32)))
    # warm up: first launches load the kernels; not part of the measurement
    for k in range(64):
        sid, p, f, n = reqs[k % n_requests]
        while not srv.submit(10 ** 7 + k, sid, p, f, n, seed=999):
            srv.poll(0)
    warm = 0
    while warm < 64:
        warm += len(srv.poll(64, wait_us=2000))
    walked0 = srv.group_walked()
    gaps = rng.exponential(1.0 / rate, n_requests)
    # scheduled arrivals in perf_counter seconds, and the offset of the
    # server's clock (both CLOCK_MONOTONIC): server_us = perf_s * 1e6 + off
    off = srv.now_us() - time.perf_counter() * 1e6
    t0 = time.perf_counter() + 1e-3
    sched = t0 + np.concatenate(([0.0], np.cumsum(gaps[:-1])))
    lat, service = np.empty(n_requests), np.empty(n_requests)
    done = 0
    i = 0
    with precise_timers() as precise:
        while done < n_requests:
            now = time.perf_counter()
            while i < n_requests and now >= sched[i]:
                sid, p, f, n = reqs[i]
                if not srv.submit(i, sid, p, f, n, seed=12345):
                    break  # queue full: the request keeps waiting, its latency keeps counting
                i += 1
            # never wait past the next arrival
            wait = 200.0 if i >= n_requests else min(200.0, max(0.0, (sched[i] - time.perf_counter()) * 1e6))
            for c in srv.poll(4096, wait_us=int(wait)):
                lat[c.tag] = c.done_us - (sched[c.tag] * 1e6 + off)
                service[c.tag] = c.done_us - c.submit_us
                done += 1
    wall = time.perf_counter() - t0
    walked = srv.group_walked() - walked0
    srv.close()
    return {"rate_req_s": rate, "requests": n_requests, "throughput_req_s": n_requests / wall,
            "group_walked_requests": walked,
            "generator_timer_slack_1ns": bool(precise),
            "p50_ms": float(np.percentile(lat, 50) / 1e3), "p99_ms": float(np.percentile(lat, 99) / 1e3),
            "p50_from_submit_ms": float(np.percentile(service, 50) / 1e3),
            "p99_from_submit_ms": float(np.percentile(service, 99) / 1e3)}


def large_model(n_sims, seed=9, device_loop=False):
    """Large polydisperse model run batched under the serving loop: 400x400x60
    um box, 5 classes (100/200/300/450/600 nm) with Dirichlet(1,...,1)
    abundances, Gaussian z detection (sigma 1 um), c = 1e10/mL (~96,000
    particles), 2000 frames, H 2000 x 1000; all requests submitted at once."""
    rng = np.random.default_rng(seed)
    m = Model(Lx=400.0, Ly=400.0, Lz=60.0, ax=160.0, ay=120.0, deltat=0.05, kmin=3, n_k=2000, n_de=1000,
              ub_de=20.0, n_species=5, detect_model=_abi.PCS_DETECT_GAUSS_Z, detect_sigma=1.0)
    D = [stokes_einstein(d) for d in (100.0, 200.0, 300.0, 450.0, 600.0)]
    srv = Server(device=0, arena_bytes=8 << 30, max_inflight=1024, device_loop=device_loop)
    sid = srv.add_system(m)
    reqs, steps = [], 0.0
    for i in range(n_sims):
        w = rng.dirichlet(np.ones(5))
        p = proposal_record(m, D, list(w * 1e10), [2.0], i, 0, 1)
        n = int(rng.poisson(1e10 * m.volume_factor))
        reqs.append((p, np.array([2000], np.uint32), np.array([n], np.uint32)))
        steps += n * 1999.0
    # warm the kernels once
    srv.submit(10 ** 6, sid, reqs[0][0], reqs[0][1], reqs[0][2], seed=1)
    while not srv.poll(4, wait_us=10000):
        pass
    # every request arrives at t0; latency runs from there (server clock)
    off = srv.now_us() - time.perf_counter() * 1e6
    t0 = time.perf_counter()
    for i, (p, f, n) in enumerate(reqs):
        while not srv.submit(i, sid, p, f, n, seed=2):
            srv.poll(0)
    lat, done = [], 0
    while done < n_sims:
        for c in srv.poll(256, wait_us=2000):
            lat.append(c.done_us - (t0 * 1e6 + off))
            done += 1
    wall = time.perf_counter() - t0
    srv.close()
    lat = np.array(lat)
    return {"sims": n_sims, "particles_per_sim": steps / n_sims / 1999.0, "sims_per_s": n_sims / wall,
            "particle_steps_per_s": steps / wall, "p50_ms": float(np.percentile(lat, 50) / 1e3),
            "p99_ms": float(np.percentile(lat, 99) / 1e3),
            # every walk launched, the warm-up request included (for profiles)
            "particle_steps_launched": steps + float(reqs[0][2][0]) * 1999.0}


def cpu_threads() -> int:
    """Host threads of this job: the CPUs it may run on, at most the 16 a
    GPU's share of the host holds on these boxes."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_workload(n=100, seed=1):
    """The CPU workload: box 200x200x30 um (periodic), window 160x120x2 um,
    H 500 x 200 over [0, 20], 1 video of 500 frames, kmin 3; n proposals from
    the prior of the reference (per component log10 c ~ U(8, 10), m ~ U(0.1,
    5) um^2/s, src/gpu/detail/simulation_impl.inl:523-537), fixed depth
    az = 2 um, Poisson particle counts (seed 1)."""
    model = Model(Lx=200.0, Ly=200.0, Lz=30.0, ax=160.0, ay=120.0, deltat=0.05, kmin=3, n_k=500, n_de=200,
                  ub_de=20.0, n_species=2)
    rng = np.random.default_rng(seed)
    jobs, props, frames, parts = [], [], [], []
    for i in range(n):
        m = sorted(rng.uniform(0.1, 5.0, 2))
        c = list(10.0 ** rng.uniform(8.0, 10.0, 2))
        p = proposal_record(model, m, c, [2.0], i, i, 1)
        props.append(p)
        frames.append(500)
        parts.append(int(rng.poisson(sum(c) * model.volume_factor)))
        jobs.append((m, c))
    return model, jobs, props, np.array(frames, np.uint32), np.array(parts, np.uint32)


def cpu_run(model, jobs, props, frames, parts, obs, threads):
    """Oracle FP64 restatement of the reference CPU path (simulate.jl +
    distance.jl), one simulation per thread; returns (sims/s, particle-steps/s, s)."""
    import ctypes as C

    import oracle_lib as O

    osys = O.system_from_abi(model.to_abi())
    work = []
    for i, (m, c) in enumerate(jobs):
        op = O.OracleProposal(props[i], frames[i:i + 1], parts[i:i + 1])
        pd = O.orc_proposal_f64()
        acc = 0.0
        for s_ in range(model.n_species):
            pd.sigma[s_] = math.sqrt(2.0 * m[s_] * model.deltat)
            pd.half_az[s_] = 1.0
            acc += c[s_] / sum(c)
            pd.cum_prob[s_] = acc
        pd.cum_prob[model.n_species - 1] = 2.0
        work.append((op, pd))
    lib = O.lib()

    def one(job):
        op, pd = job
        H = np.zeros((model.n_k, model.n_de), np.uint32)
        dr = np.zeros(1, np.uint64)
        lib.orc_simulate_f64(C.byref(osys), C.byref(op.s), C.byref(pd), model.ub_de, model.deltat, 1,
                             H.ctypes.data_as(C.c_void_p), dr.ctypes.data_as(C.c_void_p))
        return O.distance(H, obs, 0)[1]

    t0 = time.perf_counter()
    if threads == 1:
        for j in work:
            one(j)
    else:
        with cf.ThreadPoolExecutor(threads) as ex:
            list(ex.map(one, work))
    dt = time.perf_counter() - t0
    return len(work) / dt, particle_steps(frames, parts) / dt, dt


def cpu_baseline(n_one_core=25):
    """The CPU workload on all the job's host threads and on one (its first
    n_one_core proposals), and the same 100 proposals on the GPU."""
    model, jobs, props, frames, parts = cpu_workload()
    with Context(model) as c0:  # data at the truth (175/500 nm, 5e8 /mL each), seed 0
        rng = np.random.default_rng(0)
        n = int(rng.poisson(1e9 * model.volume_factor))
        p = proposal_record(model, [D175, D500], [5e8, 5e8], [2.0], 0xFFFFFFF0, 0, 1)
        b = c0.batch()
        b.load([p], np.array([500], np.uint32), np.array([n], np.uint32))
        b.run(0, _abi.PCS_MODE_BATCHED)
        _, _, H = b.read(hist=True)
        b.close()
    obs = H[0]
    threads = cpu_threads()
    sims, steps, dt = cpu_run(model, jobs, props, frames, parts, obs, threads)
    k = n_one_core
    sims1, steps1, dt1 = cpu_run(model, jobs[:k], props[:k], frames[:k], parts[:k], obs, 1)
    with Context(model, obs) as ctx:
        b = ctx.batch()
        b.load(props, frames, parts)
        walls, _, _ = run_timed(b, 5, 2, _abi.PCS_MODE_BATCHED, 300)
        b.close()
    gpu = len(props) * len(walls) / sum(walls)
    return {"value": sims, "unit": "sims/s", "cores": threads, "kind": "port",
            "particle_steps_per_s": steps, "one_core": sims1, "one_core_particle_steps_per_s": steps1,
            "host_cpus": os.cpu_count(), "gpu_same_workload_sims_per_s": gpu,
            "sample": f"CPU workload: {len(jobs)} proposals from the prior (log10 c ~ U(8,10), m ~ U(0.1,5) per "
                      f"component), 1 video x 500 frames, box 200x200x30 um, H 500x200; oracle FP64 simulate + "
                      f"dense distance, {threads} threads {dt:.1f} s; one thread on the first {k}: {dt1:.1f} s"}


def cpu_batched_sample(model, props, frames, parts, obs, n_sample, threads):
    """The same restatement on the first n_sample proposals of the batched
    workload (50/50 of 175 and 500 nm)."""
    jobs = [([D175, D500], [1.0, 1.0])] * n_sample
    sims, steps, dt = cpu_run(model, jobs, props[:n_sample], frames[:n_sample], parts[:n_sample], obs, threads)
    return {"sims_per_s": sims, "particle_steps_per_s": steps, "proposals": n_sample, "threads": threads,
            "seconds": dt}


def generation_population(model, n, rng, frames_choices=(500, 1000, 2000)):
    """Population that the batched generation proposes from: n members with
    total concentration log-uniform on 1e8..1e10 /mL split 50/50 between the
    500 nm and 175 nm species (rows m_1 <= m_2, log10 c_1, log10 c_2, depth
    2 um), equal weights, and tau = sqrt(2 var) of each row as the ABC-PMC
    driver sets it (0 for the rows without spread). Each trial has one video
    of U{500, 1000, 2000} frames. The reflected displacement keeps each
    log10 c row uniform on its prior interval, so the trials' particle counts
    span 120..12,000 as in the resident batched workload."""
    ct = 10.0 ** rng.uniform(8.0, 10.0, n)
    lc = np.log10(ct / 2.0)
    theta = np.vstack([np.full(n, D500), np.full(n, D175), lc, lc, np.full(n, 2.0)])
    w = np.full(n, 1.0 / n)
    tau = np.sqrt(2.0 * theta.var(axis=1))
    frames = rng.choice(frames_choices, n).astype(np.uint32).reshape(n, 1)
    prior = prior_record(model, (0.1, 5.0), (math.log10(0.5e8), math.log10(0.5e10)), (0.5, 4.0))
    return theta, w, tau, frames, prior


def run_generations(batch, pop, steps, warmup, seed0=1000, overlap=True):
    """Timed ABC-PMC generations on the GPU, from a population resident in
    HBM before timing (pcs_batch_set_population). One step: the proposal
    step of all the trials on the GPU (parents, displacements, Poisson
    counts, plan; the proposed parameters come back to the host), the walk
    and distance of the trials as one hipGraph launch, the read-back of the
    distances, and the trials' pmc-standard importance weights against the
    population on the GPU (n x n kernel sums over the parameters with
    spread), read back. The population is the same at every step, so every
    step does the same work. With overlap the weights are enqueued right
    after the walk (pcs_batch_pmc_weights_async) and run beside its tail and
    the distance; the step still ends only when they are back. Returns the
    wall times, per-step (propose, run + read, weights) ms and particle-steps."""
    theta, w, tau, frames, prior = pop
    n = theta.shape[1]
    batch.set_population(theta, w)
    walls, parts, psteps = [], [], []
    for i in range(warmup + steps):
        t0 = time.perf_counter()
        batch.propose(prior, None, None, tau, n, frames, seed0 + i)
        t1 = time.perf_counter()
        batch.run(seed0 + i, _abi.PCS_MODE_BATCHED)
        if overlap:
            batch.pmc_weights_async(tau)  # beside the walk: it needs the trials, not the distances
        batch.read()
        t2 = time.perf_counter()
        if overlap:
            batch.pmc_weights_wait()
        else:
            batch.pmc_weights(tau)
        t3 = time.perf_counter()
        if i >= warmup:
            walls.append(t3 - t0)
            parts.append(((t1 - t0) * 1e3, (t2 - t1) * 1e3, (t3 - t2) * 1e3))
            _, vp = batch.read_proposals()
            psteps.append(particle_steps(frames.ravel(), vp))
    return walls, parts, psteps


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batched-sims", type=int, default=100_000)
    ap.add_argument("--plain-sims", type=int, default=10_000)
    ap.add_argument("--serve-requests", type=int, default=10_000)
    ap.add_argument("--serve-rates", default="2000,5000,10000,20000,50000")
    ap.add_argument("--cpu-sample", type=int, default=384)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--large-sims", type=int, default=64)
    ap.add_argument("--z-boundary", choices=("periodic", "reflect"), default="periodic",
                    help="z walls of the 200x200x30 um box (x and y are periodic); the batched leg also runs "
                         "the other z boundary on the same proposals unless --no-other-z")
    ap.add_argument("--no-other-z", action="store_true")
    ap.add_argument("--no-weights-overlap", action="store_true",
                    help="compute the weights after the read-back instead of beside the walk")
    ap.add_argument("--tuning", default="",
                    help="launch shape 'particles_per_lane,waves_per_group,groups_per_cu[,stage_records]' instead "
                         "of the autotuner (profiling runs: no autotuner launches among the profiled ones)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="after the batched leg, copy 256 MiB device to device (pcs_selftest_copy): its known "
                         "traffic calibrates rocprofv3's FETCH_SIZE/WRITE_SIZE counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the path is sharded by proposal "
              f"ranges across GPUs by design but nothing here runs on more than one.", file=sys.stderr)
        sys.exit(2)

    model = headline_model(args.z_boundary)
    ctx = Context(model)  # scratch context to make the data histogram
    obs = observed_hist(ctx, model)
    ctx.close()
    ctx = Context(model, obs)
    log(f"observed histogram: {int(obs.sum())} tracks")

    rng = np.random.default_rng(1)
    t_pack = time.perf_counter()
    props, frames, parts = workload(model, args.batched_sims, rng, (500, 1000, 2000))
    t_pack = time.perf_counter() - t_pack
    steps_resident = particle_steps(frames, parts)
    tuning = None
    if args.tuning:
        shape = [int(x) for x in args.tuning.split(",")]
        ppl, wpg, gpc = shape[:3]
        ctx.set_tuning(ppl, wpg, gpc, shape[3] if len(shape) > 3 else 256)
        tuning = ctx.get_tuning()
        log(f"tuning: {tuning}")
    elif not args.no_autotune:
        # launch shape chosen by the autotuner on a calibration slice of the workload (untimed)
        k = min(len(props), 4000)
        tuning, tms = ctx.autotune(props[:k], frames[:k], parts[:k])
        log(f"autotune: {tuning} ({tms:.2f} ms on {k} proposals)")

    # The headline step: one ABC-PMC generation of args.batched_sims trials,
    # everything on the GPU.
    pop = generation_population(model, args.batched_sims, np.random.default_rng(11))
    gb = ctx.batch()
    walls, gparts, gsteps = run_generations(gb, pop, args.steps, args.warmup, overlap=not args.no_weights_overlap)
    reruns = gb.overflow_reruns()
    wall = sum(walls)
    sims_s = args.batched_sims * args.steps / wall
    psteps_s = sum(gsteps) / wall
    steps_per_gen = float(np.mean(gsteps))
    part_ms = np.mean(np.array(gparts), axis=0)
    # The last generation's trials again with the three batched kernels
    # launched without the graph, HIP events between them on the batch
    # stream: the walk kernel's duration.
    dwalls, total_ms, walk_ms = run_timed(gb, args.steps, 0, _abi.PCS_MODE_BATCHED_DIRECT, 100)
    _, gtracks, _ = gb.read()
    arena_cap, arena_bytes = gb.arena()
    gb.close()
    arena_use = float(gtracks.astype(np.float64).sum()) / max(arena_cap, 1)
    walk_s = float(np.mean(walk_ms)) / 1e3 if min(walk_ms) > 0 else float("nan")
    log(f"generation: {sims_s:.0f} sims/s, {psteps_s:.3e} particle-steps/s, {np.mean(walls)*1e3:.1f} ms per "
        f"generation (propose + plan {part_ms[0]:.1f} ms, walk + distance + read {part_ms[1]:.1f} ms, pmc weights "
        f"{'left after the read ' if not args.no_weights_overlap else ''}{part_ms[2]:.1f} ms), "
        f"{steps_per_gen:.3e} particle-steps per generation, walk kernel {walk_s*1e3:.1f} ms, reruns {reruns}")

    # Beside it: the same kind of generation with its proposals packed on the
    # host and resident in HBM before timing (walk + distance + read only).
    b = ctx.batch()
    t_load = time.perf_counter()
    b.load(props, frames, parts)
    t_load = time.perf_counter() - t_load
    rwalls, rtotal_ms, _ = run_timed(b, args.steps, args.warmup, _abi.PCS_MODE_BATCHED, 100)
    b.close()
    resident_s = args.batched_sims * args.steps / sum(rwalls)
    log(f"resident batch: {resident_s:.0f} sims/s ({steps_resident:.3e} particle-steps), host packing "
        f"{t_pack:.2f} s, load {t_load * 1e3:.1f} ms, graph {np.mean(rtotal_ms):.1f} ms")
    if args.calibrate_copy:
        import ctypes as C

        cms = C.c_float(0.0)
        _abi.check(_abi.load().pcs_selftest_copy(0, 256 << 20, C.byref(cms)), "pcs_selftest_copy")
        log(f"calibration: copied 268435456 bytes device to device in {cms.value:.3f} ms")

    other_z = None
    if not args.no_other_z:
        # The same proposals with the other z boundary: the reference wraps z
        # periodically (src/cpu/simulate.jl:93), the model can also reflect it.
        oz = "reflect" if args.z_boundary == "periodic" else "periodic"
        m2 = headline_model(oz)
        c2 = Context(m2)
        obs2 = observed_hist(c2, m2)
        c2.close()
        c2 = Context(m2, obs2)
        if tuning:
            c2.set_tuning(tuning["particles_per_lane"], tuning["waves_per_group"], tuning["groups_per_cu"],
                          tuning["stage_records"])
        b2 = c2.batch()
        b2.load(props, frames, parts)
        w2, _, _ = run_timed(b2, args.steps, args.warmup, _abi.PCS_MODE_BATCHED, 100)
        _, _, wk2 = run_timed(b2, args.steps, 0, _abi.PCS_MODE_BATCHED_DIRECT, 100)
        b2.close()
        c2.close()
        other_z = {"z_boundary": oz, "sims_per_s": args.batched_sims * args.steps / sum(w2),
                   "particle_steps_per_s": steps_resident * args.steps / sum(w2),
                   "walk_kernel_ms": float(np.mean(wk2))}
        log(f"batched, z {oz}: {other_z}")

    plain = None
    if not args.no_plain:
        rng2 = np.random.default_rng(2)
        pprops, pframes, pparts = workload(model, args.plain_sims, rng2, (1000,))
        psteps = particle_steps(pframes, pparts)
        pb = ctx.batch()
        pb.load(pprops, pframes, pparts)
        plain = {"sims": args.plain_sims}
        # the reference's job shape: one walk + distance pair per simulation,
        # pairs in flight on 4 streams; and the same on one stream
        for key, mode in (("streams", _abi.PCS_MODE_PLAIN_STREAMS), ("one_stream", _abi.PCS_MODE_PLAIN)):
            pw, _, _ = run_timed(pb, 1, 1, mode, 500)
            plain[key] = {"sims_per_s": args.plain_sims / pw[0], "particle_steps_per_s": psteps / pw[0],
                          "wall_s": pw[0]}
            log(f"plain ({key}): {args.plain_sims / pw[0]:.0f} sims/s over {args.plain_sims} simulations "
                f"({pw[0]:.2f} s)")
        pb.close()
        plain["sims_per_s"] = plain["streams"]["sims_per_s"]

    serve = serve_dev = None
    if not args.no_serve:
        serve, serve_dev = [], []
        for rate in [float(x) for x in args.serve_rates.split(",") if x]:
            serve.append(serving_latency(args.serve_requests, rate))
            log(f"serving (host-driven groups): {serve[-1]}")
            serve_dev.append(serving_latency(args.serve_requests, rate, device_loop=True))
            log(f"serving (device-resident loop): {serve_dev[-1]}")

    large = large_dev = None
    if args.large_sims > 0:
        large = large_model(args.large_sims)
        log(f"large polydisperse model (host-driven serving): {large}")
        large_dev = large_model(args.large_sims, device_loop=True)
        log(f"large polydisperse model (device-resident loop): {large_dev}")

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline()
        cpu["batched_workload_sample"] = cpu_batched_sample(model, props, frames, parts, obs, args.cpu_sample,
                                                            cpu["cores"])
        log(f"cpu baseline: {cpu}")

    def pick(sweep, key):
        if not sweep:
            return None
        return sweep[1][key] if len(sweep) > 1 else sweep[0][key]

    # Algorithmic work of the walk over its measured duration; issued VALU
    # lane-instructions and HBM bytes per particle-step from the committed
    # rocprofv3 runs (copy-calibrated FETCH_SIZE / WRITE_SIZE).
    prof = json.loads(PMC_PROFILE.read_text())
    ops_step = walk_ops_per_step(gauss=False)
    # the walk timed on the last generation's trials
    ach = gsteps[-1] * ops_step / walk_s if walk_s == walk_s else None
    issue = gsteps[-1] * prof["valu_lane_instr_per_particle_step"] / walk_s if walk_s == walk_s else None
    traffic = prof["hbm_bytes_per_particle_step"] * gsteps[-1]
    ratio = (sims_s / plain["sims_per_s"]) if plain else None
    ratio1 = (sims_s / plain["one_stream"]["sims_per_s"]) if plain else None
    line = {
        "metric": "ABC sims/sec & particle-steps/s, batched vs plain; p50/p99 sim latency",
        "value": sims_s,
        "unit": "sims/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+u32",
        "data": "synthetic",
        "config": {
            "workload": f"batched ABC-PMC generation on the GPU: {args.batched_sims} trials proposed from a "
                        f"resident population (total c log-uniform 1e8-1e10/mL split between 175/500 nm, "
                        f"120-12,000 particles, tau = sqrt(2 var)), frames U{{500,1000,2000}}, box 200x200x30 um "
                        f"{'periodic' if args.z_boundary == 'periodic' else 'periodic in x/y, reflecting in z'}, "
                        f"H 1000x500 uint32; step = proposal step + plan, walk + distance in one hipGraph launch, "
                        f"distances read back, pmc-standard weights of the trials against the population"
                        f"{' (enqueued beside the walk)' if not args.no_weights_overlap else ''}",
            "weights_beside_walk": not args.no_weights_overlap,
            "step_propose_ms": float(part_ms[0]),
            "step_walk_distance_read_ms": float(part_ms[1]),
            "step_weights_ms": float(part_ms[2]),
            "resident_batch_sims_per_s": resident_s,
            "resident_batch_particle_steps_per_s": steps_resident * args.steps / sum(rwalls),
            "z_boundary": args.z_boundary,
            "particle_steps_per_generation": steps_per_gen,
            "particle_steps_per_s": psteps_s,
            # scalar copies of the nested figures below
            "other_z_boundary_sims_per_s": other_z["sims_per_s"] if other_z else None,
            "other_z_boundary_walk_ms": other_z["walk_kernel_ms"] if other_z else None,
            "plain_sims_per_s": plain["sims_per_s"] if plain else None,
            "plain_one_stream_sims_per_s": plain["one_stream"]["sims_per_s"] if plain else None,
            "batched_vs_plain": ratio,
            "batched_vs_plain_one_stream": ratio1,
            "resident_vs_plain": (resident_s / plain["sims_per_s"]) if plain else None,
            # headline latency: the device-resident loop at the second swept rate (5,000 req/s by
            # default), from each request's scheduled arrival
            "p50_ms": pick(serve_dev, "p50_ms"),
            "p99_ms": pick(serve_dev, "p99_ms"),
            "latency_form": ("device-resident loop (requests of <= 1 tile per CU admitted behind <= 12 "
                             "unfinished ones walked a group per tile), from scheduled arrival") if serve_dev else None,
            "host_driven_p50_ms": pick(serve, "p50_ms"),
            "host_driven_p99_ms": pick(serve, "p99_ms"),
            "large_polydisperse_sims_per_s": large_dev["sims_per_s"] if large_dev else None,
            "large_polydisperse_particle_steps_per_s": large_dev["particle_steps_per_s"] if large_dev else None,
            "cpu_one_core_sims_per_s": cpu["one_core"] if cpu else None,
            "overflow_reruns": reruns,
            "host_packing_s": t_pack,
            "load_ms": t_load * 1e3,
            "arena_bytes": arena_bytes,
            "arena_use": arena_use,
            "resident_graph_device_ms": float(np.mean(rtotal_ms)),
            "direct_launch_wall_ms": float(np.mean(dwalls)) * 1e3,
            "parallelism": "single GPU",
            "tuning": tuning,
            "other_z_boundary": other_z,
            "plain": plain,
            "serving": serve,
            "serving_device_loop": serve_dev,
            "large_polydisperse": large,
            "large_polydisperse_device_loop": large_dev,
        },
        "roofline": {
            "bound": "valu",
            "achieved": ach / 1e12 if ach else None,
            "peak": FP32_PEAK / 1e12,
            "unit": "TFLOP/s",
            "frac": (ach / FP32_PEAK) if ach else None,
            "traffic": traffic,
            "kernel": "walk_kernel",
            "kernel_ms": walk_s * 1e3,
            "algorithmic_ops_per_particle_step": ops_step,
            # the same fraction with one Philox word per unit normal counted
            # (0.75 blocks per step: 123 ops), the count this stream's two
            # blocks per four steps undercut
            "frac_one_word_per_normal": (ach / FP32_PEAK * walk_ops_per_step(words_per_normal=1.0) / ops_step)
            if ach else None,
            "issue_slot_util": (issue / ISSUE_PEAK) if issue else None,
            "issued_valu_lane_instr_per_particle_step": prof["valu_lane_instr_per_particle_step"],
            "issued_and_traffic_from": str(PMC_PROFILE.relative_to(ROOT)),
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
