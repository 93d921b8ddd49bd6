"""CPU ABC-PMC restatement of the reference's Julia estimator
(src/cpu/estimate.jl:1-240) driving the oracle's FP64 simulate
(orc_simulate_f64, restating src/cpu/simulate.jl) and the exact distance
(orc_distance, restating src/cpu/distance.jl). Test infrastructure: it is
the CPU side of the posterior comparison, never part of the product.

Where the reference's two estimators differ this follows the Julia one: a
fixed depth is C identical component rows, each displaced member keeps them
equal, and the pmc-standard weight multiplies the depth kernel once per row
(estimate.jl:60-61, 143-144, 199-203; the C++ driver multiplies it once,
src/gpu/detail/simulation_impl.inl:849-852); the random member is the first
whose cumulative weight reaches u (rand_weighted_index.jl); displacement
reflects into the open interval (displace.jl:5).

Differences from the Julia code, all deliberate: the observed and simulated
histograms bin with trunc (as the GPU reference) instead of ceil; the
adaptive-gamma pass simulates 10^log10c (estimate.jl:87-100 passes log10c);
the random streams are numpy's and the oracle's Philox; the simulation is
the oracle's FP64 walk (orc_simulate_f64) rather than simulate.jl's own
stream.
"""
from __future__ import annotations

import concurrent.futures as cf
import ctypes as C
import math
import threading

import numpy as np

import oracle_lib as O
from part_conc_spt_abc_amd.model import Model, proposal_record


def normpdf(x, mu, sigma):  # src/cpu/normpdf.jl
    return 1.0 / (math.sqrt(2.0 * math.pi) * sigma) * np.exp(-0.5 * (x - mu) ** 2 / sigma ** 2)


def displace(rng, x, tau, lb, ub):  # src/cpu/displace.jl (open bounds)
    xb = x + tau * rng.standard_normal()
    while not (lb < xb < ub):
        if xb < lb:
            xb = xb + 2.0 * (lb - xb)
        if xb > ub:
            xb = xb - 2.0 * (xb - ub)
    return xb


class CpuAbc:
    def __init__(self, model: Model, frames, K, DE, lb_m, ub_m, lb_c, ub_c, lb_az, ub_az, n_samples, fixed_depth=True,
                 weighting="pmc-standard", seed=0, threads=8):
        self.model = model
        self.frames = np.asarray(frames, np.uint32)
        self.osys = O.system_from_abi(model.to_abi())
        self.H, _, _ = O.reference_hist(K, DE, model.ub_de, model.n_k, model.n_de)
        self.b = dict(m=(lb_m, ub_m), c=(math.log10(lb_c), math.log10(ub_c)), az=(lb_az, ub_az))
        self.N = n_samples
        self.C = model.n_species
        self.fixed = fixed_depth
        self.weighting = weighting
        self.rng = np.random.default_rng(seed)
        self.threads = threads
        self.sim_counter = 0
        self.lock = threading.Lock()

    def simulate(self, m, log10c, az, rng, seed):
        """One trial: Poisson particle counts per video, FP64 walk, exact distance."""
        c = 10.0 ** np.asarray(log10c)
        lam = c.sum() * self.model.volume_factor
        parts = rng.poisson(lam, len(self.frames)).astype(np.uint32)
        p = proposal_record(self.model, list(m), list(c), list(az), 0, 0, len(self.frames))
        op = O.OracleProposal(p, self.frames, parts)
        pd = O.orc_proposal_f64()
        acc = 0.0
        for i in range(self.C):
            pd.sigma[i] = math.sqrt(2.0 * m[i] * self.model.deltat)
            pd.half_az[i] = 0.5 * (az[i] if len(az) > 1 else az[0])
            acc += c[i] / c.sum()
            pd.cum_prob[i] = acc
        pd.cum_prob[self.C - 1] = 2.0
        Hs, _, _ = O.simulate_f64(self.osys, op, pd, self.model.ub_de, self.model.deltat, seed)
        return O.distance(Hs, self.H, 0)[1]

    def run(self, gamma_initial, delta_gamma, generations, ub_trials=10 ** 9):
        N, C = self.N, self.C
        lb, ub = self.b["m"]
        m = lb + (ub - lb) * self.rng.random((C, N))
        lb, ub = self.b["c"]
        lc = lb + (ub - lb) * self.rng.random((C, N))
        lb, ub = self.b["az"]
        az = lb + (ub - lb) * (np.repeat(self.rng.random((1, N)), C, 0) if self.fixed else self.rng.random((C, N)))
        w = np.full(N, 1.0 / N)
        tau = lambda x: np.sqrt(2.0 * x.var(axis=1))  # noqa: E731  var(corrected=false)
        tm, tc, ta = tau(m), tau(lc), tau(az)
        gamma = gamma_initial
        dist = np.zeros(N)
        sims = 0
        for gen in range(generations):
            gamma -= delta_gamma
            cum_w = np.cumsum(w)
            trial_count = np.zeros(N, np.int64)
            m_s, lc_s, az_s, d_s = m.copy(), lc.copy(), az.copy(), np.zeros(N)
            seeds = self.rng.integers(0, 2 ** 63, size=N)

            def member(j):
                rng = np.random.default_rng(seeds[j])
                d = math.inf
                while d > gamma and trial_count.mean() < ub_trials:
                    idx = int(np.searchsorted(cum_w, rng.random()))  # rand_weighted_index: first cum >= u
                    idx = min(idx, N - 1)
                    mb = np.array([displace(rng, m[i, idx], tm[i], *self.b["m"]) for i in range(C)])
                    cb = np.array([displace(rng, lc[i, idx], tc[i], *self.b["c"]) for i in range(C)])
                    if self.fixed:
                        a0 = displace(rng, az[0, idx], ta[0], *self.b["az"])
                        ab = np.full(C, a0)
                    else:
                        ab = np.array([displace(rng, az[i, idx], ta[i], *self.b["az"]) for i in range(C)])
                    p = np.argsort(mb, kind="stable")
                    mb, cb, ab = mb[p], cb[p], ab[p]
                    with self.lock:
                        self.sim_counter += 1
                        sseed = self.sim_counter
                    d = self.simulate(mb, cb, ab, rng, sseed)
                    with self.lock:
                        trial_count[j] += 1
                m_s[:, j], lc_s[:, j], az_s[:, j], d_s[j] = mb, cb, ab, d

            with cf.ThreadPoolExecutor(self.threads) as ex:
                list(ex.map(member, range(N)))
            sims += int(trial_count.sum())
            if trial_count.mean() >= ub_trials:
                gamma += delta_gamma
                break
            if self.weighting == "pmc-standard":
                ws = np.zeros(N)
                for j in range(N):
                    term = np.ones(N)
                    for i in range(C):
                        term *= normpdf(m_s[i, j] - m[i], 0.0, tm[i])
                        term *= normpdf(lc_s[i, j] - lc[i], 0.0, tc[i])
                    # one depth kernel per component row, also for a fixed
                    # depth, whose C rows are equal (estimate.jl:199-203)
                    for i in range(C):
                        term *= normpdf(az_s[i, j] - az[i], 0.0, ta[i])
                    ws[j] = 1.0 / (w * term).sum()
                w = ws / ws.sum()
            else:
                w = 1.0 / d_s ** 2
                w = w / w.sum()
            m, lc, az, dist = m_s, lc_s, az_s, d_s
            tm, tc, ta = tau(m), tau(lc), tau(az)
        return dict(m=m, c=10.0 ** lc, az=az, dist=dist, w=w, gamma=gamma, simulations=sims)
