"""The ABC-PMC proposal step on the GPU (pcs_batch_propose) against the
oracle's restatement (orc_propose): proposed parameters, proposal records
(half depths, cumulative abundances, step sizes) and Poisson particle counts
bit for bit, for variable and fixed depth; the simulations of a proposed
batch equal the oracle's on the same records (graph and plain forms, with a
repeat of the step on the same batch, and with record slices forced to
overflow); the displacement of a single-member population is N(0, tau^2)."""
import math

import numpy as np
import pytest
from scipy import stats

import cases
import oracle_lib as O
from part_conc_spt_abc_amd import _abi
from part_conc_spt_abc_amd.gpu import Context
from part_conc_spt_abc_amd.model import prior_record

pytestmark = pytest.mark.gpu

M_B, C_B, AZ_B = (0.1, 5.0), (8.0, 10.0), (0.5, 4.0)


def population(rng, C, Z, n_pop):
    theta = np.vstack([rng.uniform(*M_B, (C, n_pop)), rng.uniform(8.0, 9.6, (C, n_pop)),
                       rng.uniform(*AZ_B, (Z, n_pop))])
    w = rng.dirichlet(np.ones(n_pop))
    tau = np.sqrt(2.0 * theta.var(axis=1))
    return theta, np.cumsum(w), tau


def oracle_prior(p):
    return O.orc_prior(p.n_components, p.n_depths, p.m_lo, p.m_hi, p.log10c_lo, p.log10c_hi, p.az_lo, p.az_hi,
                       p.deltat, p.volume)


def device_propose(ctx, pr, theta, cum, tau, n, frames, seed, first=0):
    b = ctx.batch()
    th = b.propose(pr, theta, cum, tau, n, frames, seed, first_sim_id=first)
    recs, parts = b.read_proposals()
    return b, th, recs, parts


@pytest.mark.parametrize("C,fixed", [(3, False), (2, True)])
def test_device_proposals_equal_the_oracle(gpu_lib, C, fixed):
    model = cases.headline_model(n_k=500, n_de=200, n_species=C)
    Z = 1 if fixed else C
    rng = np.random.default_rng(10 + C)
    theta, cum, tau = population(rng, C, Z, 700)
    n, J = 3000, 2
    frames = rng.choice([200, 500], (n, J)).astype(np.uint32)
    pr = prior_record(model, M_B, C_B, AZ_B, fixed_depth=fixed)
    with Context(model) as ctx:
        b, th, recs, parts = device_propose(ctx, pr, theta, cum, tau, n, frames, 0xABCDEF12345, first=100)
        b.close()
    o = O.propose(oracle_prior(pr), theta, cum, tau, n, J, 0xABCDEF12345)
    assert np.array_equal(th.view(np.uint64), o["theta"].view(np.uint64))
    assert np.array_equal(parts.reshape(n, J), o["particles"])
    for t in (0, 1, n // 2, n - 1):
        assert recs[t].sim_id == 100 + t and recs[t].first_video == t * J and recs[t].n_videos == J
    f32 = lambda recs, f: np.array([list(getattr(r, f)) for r in recs], np.float32)  # noqa: E731
    for f in ("half_az", "cum_prob", "sigma"):
        assert np.array_equal(f32(recs, f).view(np.uint32), o[f].view(np.uint32)), f
    print(f"C={C} fixed={fixed}: mean particles {parts.mean():.1f}, parents {len(np.unique(o['parent']))}",
          flush=True)


def test_proposed_batch_simulations_equal_the_oracle(gpu_lib):
    """Proposal step, then the walk and distance of the same batch: graph
    form, plain form, and the step repeated on the batch with a new seed
    (the captured graph and the arena are reused); a sample of trials equals
    the oracle's simulation of the oracle's own record."""
    model = cases.cpu_model()  # H 500 x 200
    rng = np.random.default_rng(3)
    obs = np.zeros((model.n_k, model.n_de), np.uint32)
    obs[2:40, :60] = rng.integers(0, 4, (38, 60))
    theta, cum, tau = population(rng, 2, 1, 300)
    n, J = 400, 1
    frames = rng.choice([300, 500], (n, J)).astype(np.uint32)
    pr = prior_record(model, M_B, C_B, AZ_B)
    osys = O.system_from_abi(model.to_abi())
    with Context(model, obs) as ctx:
        b = ctx.batch()
        for step, pseed in enumerate((21, 22)):
            b.propose(pr, theta, cum, tau, n, frames, pseed)
            recs, parts = b.read_proposals()
            got = {}
            for mode in (_abi.PCS_MODE_BATCHED, _abi.PCS_MODE_PLAIN):
                b.run(77, mode)
                dist, tracks, _ = b.read()
                got[mode] = (dist.copy(), tracks.copy(), b.read_exact().copy())
            assert np.array_equal(got[_abi.PCS_MODE_BATCHED][2], got[_abi.PCS_MODE_PLAIN][2])
            assert np.array_equal(got[_abi.PCS_MODE_BATCHED][1], got[_abi.PCS_MODE_PLAIN][1])
            dist, tracks, exact = got[_abi.PCS_MODE_BATCHED]
            pick = list(np.argsort(parts)[:3]) + list(np.argsort(parts)[-3:]) + list(rng.choice(n, 4, replace=False))
            o = O.propose(oracle_prior(pr), theta, cum, tau, n, J, pseed)
            for t in pick:
                t = int(t)
                # the oracle's own record of trial t
                rec = _abi.pcs_proposal(t, t * J, J, 0, tuple(o["half_az"][t]), tuple(o["cum_prob"][t]),
                                        tuple(o["sigma"][t]))
                H, total, dropped = O.simulate_f32(osys, O.OracleProposal(rec, frames[t], o["particles"][t]), 77)
                ex, lg = O.distance(H, obs, 0)
                assert (int(exact[t]), int(tracks[t])) == (ex, total - dropped), (step, t)
                assert dist[t] == pytest.approx(lg, rel=0, abs=1e-12)
        b.close()


def test_proposed_batch_overflow_reruns(gpu_lib):
    model = cases.cpu_model()
    rng = np.random.default_rng(4)
    theta, cum, tau = population(rng, 2, 1, 100)
    n = 60
    frames = np.full((n, 1), 400, np.uint32)
    pr = prior_record(model, M_B, C_B, AZ_B)
    with Context(model) as ctx:
        b = ctx.batch()
        b.propose(pr, theta, cum, tau, n, frames, 5)
        b.run(9, _abi.PCS_MODE_BATCHED)
        d0, t0, _ = b.read()
        e0 = b.read_exact()
        b.close()
        ctx.set_record_capacity_factor(1e-3)
        b = ctx.batch()
        b.propose(pr, theta, cum, tau, n, frames, 5)
        b.run(9, _abi.PCS_MODE_BATCHED)
        d1, t1, _ = b.read()
        e1 = b.read_exact()
        reruns = b.overflow_reruns()
        # the batch runs again after its reruns: the device plan is restored
        b.run(9, _abi.PCS_MODE_BATCHED)
        d2, t2, _ = b.read()
        b.close()
    assert reruns > 0
    assert np.array_equal(t0, t1) and np.array_equal(e0, e1) and np.array_equal(d0.view(np.uint64), d1.view(np.uint64))
    assert np.array_equal(t0, t2) and np.array_equal(d0.view(np.uint64), d2.view(np.uint64))


def test_displacement_is_gaussian(gpu_lib):
    """A one-member population, bounds far away: proposed - member over tau
    is N(0, 1) per parameter (KS and moments)."""
    model = cases.headline_model(n_k=50, n_de=20, n_species=1)
    theta = np.array([[2.0], [8.5], [2.0]])
    tau = np.array([0.05, 0.02, 0.1])
    pr = prior_record(model, (0.0, 1e3), (-20.0, 15.0), (0.0, 1e3))
    pr.volume = 1e-12  # about no particles: only the parameters matter here
    n = 20000
    with Context(model) as ctx:
        b = ctx.batch()
        th = b.propose(pr, theta, np.array([1.0]), tau, n, np.full((n, 1), 2, np.uint32), 31)
        b.close()
    z = (th - theta) / tau[:, None]
    for p in range(3):
        assert stats.kstest(z[p], "norm").pvalue > 1e-4, p
        assert abs(z[p].mean()) < 5 / math.sqrt(n) and abs(z[p].var() - 1) < 5 * math.sqrt(2 / n)


def test_resident_population_equals_host_arrays(gpu_lib):
    """A population set resident in the batch (cumulative weights formed on
    the device) gives the same proposals as the host arrays with numpy's
    cumsum, and pcs_batch_pmc_weights equals pcs_pmc_weights on the same
    values (rows with tau 0 left out), bit for bit."""
    from part_conc_spt_abc_amd.gpu import pmc_weights

    model = cases.headline_model(n_k=100, n_de=50, n_species=2)
    rng = np.random.default_rng(17)
    theta, _, tau = population(rng, 2, 1, 3000)
    theta[4] = 2.0  # a depth without spread: tau 0, left out of the weights
    tau = np.sqrt(2.0 * theta.var(axis=1))
    assert tau[4] == 0.0
    w = rng.dirichlet(np.ones(theta.shape[1]))
    n = 2500
    frames = np.full((n, 1), 40, np.uint32)
    pr = prior_record(model, M_B, C_B, AZ_B)
    with Context(model) as ctx:
        b = ctx.batch()
        th_host = b.propose(pr, theta, np.cumsum(w), tau, n, frames, 99)
        _, parts_host = b.read_proposals()
        b.set_population(theta, w)
        th_res = b.propose(pr, None, None, tau, n, frames, 99)
        _, parts_res = b.read_proposals()
        w_res = b.pmc_weights(tau)
        b.close()
    assert np.array_equal(th_host.view(np.uint64), th_res.view(np.uint64))
    assert np.array_equal(parts_host, parts_res)
    keep = tau > 0
    w_ref = pmc_weights(th_res[keep], theta[keep], w, tau[keep])
    assert np.array_equal(w_res.view(np.uint64), w_ref.view(np.uint64))
    assert abs(w_res.sum() - 1.0) < 1e-12


def test_weights_beside_the_walk_equal_the_synchronous_ones(gpu_lib):
    """pcs_batch_pmc_weights_async enqueued after the walk of the same
    trials (so the two run side by side) and collected after the read-back
    gives the synchronous weights bit for bit, and leaves the walk's results
    unchanged; _wait without a pending computation, or after a new proposal
    step dropped it, returns PCS_ESTATE."""
    model = cases.headline_model(n_k=300, n_de=100, n_species=2)
    rng = np.random.default_rng(23)
    theta, _, tau = population(rng, 2, 1, 5000)
    w = rng.dirichlet(np.ones(theta.shape[1]))
    n = 6000
    frames = np.full((n, 1), 300, np.uint32)
    pr = prior_record(model, M_B, C_B, AZ_B)
    with Context(model) as ctx:
        b = ctx.batch()
        b.set_population(theta, w)
        b.propose(pr, None, None, tau, n, frames, 7)
        b.run(3, _abi.PCS_MODE_BATCHED)
        d_sync, t_sync, _ = b.read()
        w_sync = b.pmc_weights(tau)
        b.run(3, _abi.PCS_MODE_BATCHED)
        b.pmc_weights_async(tau)
        d_async, t_async, _ = b.read()
        w_async = b.pmc_weights_wait()
        with pytest.raises(_abi.PcsError) as e:
            b.pmc_weights_wait()
        assert e.value.code == _abi.PCS_ESTATE
        b.pmc_weights_async(tau)
        b.propose(pr, None, None, tau, n, frames, 8)
        with pytest.raises(_abi.PcsError) as e:
            b.pmc_weights_wait()
        assert e.value.code == _abi.PCS_ESTATE
        b.close()
    assert np.array_equal(w_async.view(np.uint64), w_sync.view(np.uint64))
    assert np.array_equal(t_async, t_sync) and np.array_equal(d_async.view(np.uint64), d_sync.view(np.uint64))
