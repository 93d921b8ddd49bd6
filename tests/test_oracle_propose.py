"""Pins the oracle's ABC-PMC proposal step (orc_propose and its parts).

- exp10: mpmath's 10^x, within 2 ulp on a grid over [-30, 30].
- Poisson counts: the pmf of scipy.stats.poisson by chi-square for small
  lambda, mean and variance for large lambda, and 0 at lambda = 0.
- parents: frequencies against the weights (chi-square); the first member
  whose cumulative weight reaches u (a weight-0 member is never picked).
- displacement: (proposed - parent) / tau ~ N(0, 1) (KS, moments, no
  correlation between parameters) away from the bounds; a population
  uniform on the prior interval stays uniform after the reflected step
  (the reflection kernel preserves it); tau = 0 copies the parent.
- per-trial precompute: components sorted by m with their c and az, cum_prob
  increasing to the forced 2, sigma = sqrt(2 m dt), half_az = az / 2, and
  particle counts with mean lambda = sum c * volume.
"""
import math

import mpmath as mp
import numpy as np
import pytest
from scipy import stats

import oracle_lib as O


def ulp_err(a, exact):
    a = float(a)
    return abs(mp.mpf(a) - exact) / mp.mpf(math.ulp(float(exact)))


def test_exp10_within_two_ulp():
    mp.mp.prec = 120
    xs = list(np.linspace(-30.0, 30.0, 6001)) + list(np.random.default_rng(0).uniform(6.0, 11.0, 4000))
    worst = max(ulp_err(O.exp10(x), mp.power(10, mp.mpf(float(x)))) for x in xs)
    assert worst <= 2.0, worst
    assert O.exp10(0.0) == 1.0


@pytest.mark.parametrize("lam", [0.3, 2.5, 17.0, 140.0])
def test_poisson_pmf_small_lambda(lam):
    n = 20000
    k = np.array([O.poisson(lam, 123, t, 1) for t in range(n)])
    hi = int(stats.poisson.ppf(1 - 1e-6, lam)) + 1
    lo = int(stats.poisson.ppf(1e-6, lam))
    obs = np.array([np.sum(k <= lo)] + [np.sum(k == v) for v in range(lo + 1, hi)] + [np.sum(k >= hi)], float)
    exp = n * np.array([stats.poisson.cdf(lo, lam)] + [stats.poisson.pmf(v, lam) for v in range(lo + 1, hi)]
                       + [stats.poisson.sf(hi - 1, lam)])
    keep = exp >= 5
    chi2 = float(((obs[keep] - exp[keep]) ** 2 / exp[keep]).sum() + 0.0)
    dof = int(keep.sum()) - 1
    assert stats.chi2.sf(chi2, dof) > 1e-4, (chi2, dof)


@pytest.mark.parametrize("lam", [1200.0, 12000.0])
def test_poisson_moments_large_lambda(lam):
    n = 1500
    k = np.array([O.poisson(lam, 77, t, 3) for t in range(n)], float)
    se_mean = math.sqrt(lam / n)
    assert abs(k.mean() - lam) < 5 * se_mean
    se_var = lam * math.sqrt(2.0 / (n - 1))
    assert abs(k.var(ddof=1) - lam) < 5 * se_var


def test_poisson_zero():
    assert all(O.poisson(0.0, 5, t, 0) == 0 for t in range(50))


def prior(C, Z, m=(0.1, 5.0), c=(8.0, 10.0), az=(0.5, 4.0), deltat=0.05, volume=1.2e-6):
    return O.orc_prior(C, Z, m[0], m[1], c[0], c[1], az[0], az[1], deltat, volume)


def test_parents_follow_the_weights():
    rng = np.random.default_rng(1)
    n_pop = 40
    w = rng.dirichlet(np.ones(n_pop))
    w[7] = 0.0
    w /= w.sum()
    cum = np.cumsum(w)
    theta = np.vstack([rng.uniform(1, 2, n_pop), rng.uniform(8.5, 9.5, n_pop), rng.uniform(1, 3, n_pop)])
    out = O.propose(prior(1, 1), theta, cum, np.zeros(3), 60000, 1, 9)
    counts = np.bincount(out["parent"], minlength=n_pop)
    assert counts[7] == 0
    keep = w > 0
    exp = 60000 * w[keep]
    chi2 = float(((counts[keep] - exp) ** 2 / exp).sum())
    assert stats.chi2.sf(chi2, int(keep.sum()) - 1) > 1e-4
    # tau = 0: every trial is its parent's parameters
    assert np.array_equal(out["theta"], theta[:, out["parent"]])


def test_displacement_is_gaussian_away_from_the_bounds():
    rng = np.random.default_rng(2)
    n_pop = 50
    theta = np.vstack([rng.uniform(2, 3, n_pop), rng.uniform(8.8, 9.2, n_pop), rng.uniform(1.5, 2.5, n_pop)])
    tau = np.array([0.05, 0.02, 0.03])
    pr = prior(1, 1, m=(-1e6, 1e6), c=(-1e6, 1e6), az=(-1e6, 1e6))
    out = O.propose(pr, theta, np.cumsum(np.full(n_pop, 1.0 / n_pop)), tau, 40000, 1, 11)
    z = (out["theta"] - theta[:, out["parent"]]) / tau[:, None]
    for p in range(3):
        assert stats.kstest(z[p], "norm").pvalue > 1e-4, p
        assert abs(z[p].mean()) < 5 / math.sqrt(z.shape[1])
        assert abs(z[p].var() - 1.0) < 5 * math.sqrt(2.0 / z.shape[1])
    c = np.corrcoef(z)
    assert np.all(np.abs(c[np.triu_indices(3, 1)]) < 5 / math.sqrt(z.shape[1]))


def test_reflection_keeps_a_uniform_population_uniform():
    rng = np.random.default_rng(3)
    n_pop = 4000
    lo, hi = 8.0, 10.0
    theta = np.vstack([rng.uniform(0.1, 5, n_pop), rng.uniform(lo, hi, n_pop), rng.uniform(0.5, 4, n_pop)])
    tau = np.array([0.3, math.sqrt(2 * (hi - lo) ** 2 / 12), 0.8])  # tau of the driver, sqrt(2 var)
    out = O.propose(prior(1, 1, c=(lo, hi)), theta, np.cumsum(np.full(n_pop, 1.0 / n_pop)), tau, 40000, 1, 12)
    y = out["theta"][1]
    assert y.min() >= lo and y.max() <= hi
    assert stats.kstest((y - lo) / (hi - lo), "uniform").pvalue > 1e-4
    # and the other rows stay inside their intervals
    assert out["theta"][0].min() >= 0.1 and out["theta"][0].max() <= 5.0
    assert out["theta"][2].min() >= 0.5 and out["theta"][2].max() <= 4.0


@pytest.mark.parametrize("C,Z", [(3, 3), (2, 1), (5, 1)])
def test_sorted_components_and_precompute(C, Z):
    rng = np.random.default_rng(4 + C)
    n_pop = 64
    P = 2 * C + Z
    theta = np.vstack([rng.uniform(0.1, 5, (C, n_pop)), rng.uniform(8, 10, (C, n_pop)), rng.uniform(0.5, 4, (Z, n_pop))])
    tau = np.zeros(P)
    pr = prior(C, Z, deltat=0.05, volume=1.2e-6)
    out = O.propose(pr, theta, np.cumsum(np.full(n_pop, 1.0 / n_pop)), tau, 3000, 2, 13)
    par = theta[:, out["parent"]]
    order = np.argsort(par[:C], axis=0, kind="stable")
    cols = np.arange(par.shape[1])
    assert np.array_equal(out["theta"][:C], par[:C][order, cols])
    assert np.array_equal(out["theta"][C:2 * C], par[C:2 * C][order, cols])
    if Z > 1:
        assert np.array_equal(out["theta"][2 * C:], par[2 * C:][order, cols])
    else:
        assert np.array_equal(out["theta"][2 * C], par[2 * C])
    m, lc = out["theta"][:C], out["theta"][C:2 * C]
    c = 10.0 ** lc
    frac = np.cumsum(c, axis=0) / c.sum(axis=0)
    cp = out["cum_prob"].T
    assert np.all(cp[C - 1] == 2.0) and np.all(cp[C:] == 0.0)
    assert np.all(np.diff(cp[:C - 1], axis=0) >= 0)
    assert np.allclose(cp[:C - 1], frac[:C - 1], rtol=1e-6, atol=1e-7)
    sg = out["sigma"].T
    assert np.allclose(sg[:C], np.sqrt(2 * m * 0.05), rtol=2e-7) and np.all(sg[C:] == 0.0)
    hz = out["half_az"].T
    az = out["theta"][2 * C:] if Z > 1 else np.repeat(out["theta"][2 * C][None], C, 0)
    assert np.array_equal(hz[:C], (0.5 * az.astype(np.float32)).astype(np.float32)) and np.all(hz[C:] == 0.0)
    lam = c.sum(axis=0) * 1.2e-6
    k = out["particles"].astype(float)
    z = (k - lam[:, None]).sum() / math.sqrt(2 * lam.sum())
    assert abs(z) < 5.0
