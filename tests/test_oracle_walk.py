"""Pins the oracle's walk / detect / track / bin restatement
(src/cpu/simulate.jl:56-124, src/gpu/cusim/kernel-simulate-unroll.inl:99-295)
with closed forms, invariants and brute force:

- sigma = 0: every particle that starts inside is one track of length F
  with DE = 0 (exact), and the starting fraction is the volume ratio;
- detection everywhere: exactly one track of length F per particle and
  DE ~ D chi2_{2(F-1)} / (2(F-1)) (periodic wraps do not change dx);
- trajectories: positions stay in the box, wrapped increments are
  N(0, sigma^2) and uncorrelated, reflecting walls keep z inside;
- tracks re-derived from the trajectory (runs of detected frames, first
  point excluded from the MSD, kmin, flush at the end) equal the oracle's;
- the FP64 twin (libm math, FP64 positions) gives the same histogram up to
  rare boundary cases; FP32 position drift is bounded;
- reflecting and periodic z give histograms of the same law (centred window).
"""
import math

import numpy as np
import pytest
from scipy import stats

import cases
import oracle_lib as O
from part_conc_spt_abc_amd.model import Model, proposal_record


def small_model(**kw):
    args = dict(Lx=200.0, Ly=200.0, Lz=30.0, ax=160.0, ay=120.0, deltat=0.05, kmin=3, n_k=100, n_de=200,
                ub_de=20.0, n_species=2)
    args.update(kw)
    return Model(**args)


def run(model, m, c, az, N, F, seed=1, sim_id=3):
    p = proposal_record(model, m, c, az, sim_id, 0, len(F))
    osys = O.system_from_abi(model.to_abi())
    op = O.OracleProposal(p, F, N)
    return osys, op, p


def test_sigma_zero_tracks_are_exact():
    model = small_model(n_k=50)
    N, F = 20000, 50
    osys, op, p = run(model, [0.0, 0.0], [5e8, 5e8], [2.0, 6.0], [N], [F])
    H, total, dropped = O.simulate_f32(osys, op, 9)
    # independent count of particles starting inside, from the block-0 words
    inside = 0
    h = [np.float32(100.0), np.float32(100.0), np.float32(15.0)]
    for i in range(N):
        w = O.philox((0, i, 3, 0), (9, 0))  # counter (video 0, particle i, sim 3, block 0)
        u = (w >> 8).astype(np.float32) * np.float32(2.0 ** -24)
        x, y, z = (h[a] * (np.float32(2.0) * u[a] - np.float32(1.0)) for a in range(3))
        s = 0 if u[3] < p.cum_prob[0] else 1
        haz = np.float32(1.0) if s == 0 else np.float32(3.0)
        inside += abs(z) <= haz and abs(x) <= np.float32(80.0) and abs(y) <= np.float32(60.0)
    assert dropped == 0
    assert total == inside == H.sum() == H[F - 1, 0]
    # uniform start: fraction = (160*120/200^2) * (0.5*2/30 + 0.5*6/30)
    f = 0.48 * (0.5 * 2 / 30 + 0.5 * 6 / 30)
    assert abs(inside - N * f) < 5 * math.sqrt(N * f * (1 - f))


@pytest.mark.parametrize("F", [2, 6, 25])
def test_full_detection_de_law(F):
    D = 2.5
    model = small_model(ax=200.0, ay=200.0, n_species=1, kmin=2, n_k=64, n_de=4000, ub_de=40.0)
    N = 20000
    osys, op, p = run(model, [D], [1e9], [60.0], [N], [F])
    K, DE = O.tracks_f32(osys, op, 5)
    assert len(K) == N and np.all(K == F)
    dof = 2 * (F - 1)
    x = DE.astype(np.float64) * dof / D
    assert stats.kstest(x, stats.chi2(dof).cdf).pvalue > 1e-4
    assert abs(DE.mean() - D) < 5 * D * math.sqrt(2.0 / dof / N)


@pytest.mark.parametrize("zref", [0, 1])
def test_trajectory_increments_and_walls(zref):
    model = small_model(Lx=2.0, Ly=2.0, Lz=2.0, ax=1.0, ay=1.0, z_boundary=zref)
    sigma2 = 2 * 0.9 * 0.05
    osys, op, p = run(model, [0.9, 0.9], [1e9, 1e9], [1.0], [4], [2000])
    pd = O.orc_proposal_f64()
    x32, x64 = O.trace_particle(osys, op, pd, 2, 0, 1)
    assert np.all(np.abs(x32) <= 1.0)
    d = np.diff(x32[:, :2].astype(np.float64), axis=0)
    d = (d + 1.0) % 2.0 - 1.0  # undo single periodic wraps (|step| << L)
    n = d.size
    assert abs(d.var() / sigma2 - 1.0) < 5 * math.sqrt(2.0 / n)
    a = d[:, 0]
    r1 = np.corrcoef(a[:-1], a[1:])[0, 1]
    assert abs(r1) < 5 / math.sqrt(len(a))


def tracks_from_trajectory(xyz, hax, hay, haz, kmin, four_dt):
    det = (np.abs(xyz[:, 0]) <= hax) & (np.abs(xyz[:, 1]) <= hay) & (np.abs(xyz[:, 2]) <= haz)
    out = []
    k, de = 0, 0.0
    for t in range(len(xyz)):
        if det[t]:
            k += 1
            if k >= 2:
                d = xyz[t, :2].astype(np.float64) - xyz[t - 1, :2].astype(np.float64)
                d = (d + 100.0) % 200.0 - 100.0
                de += float((d ** 2).sum())
        else:
            if k >= kmin:
                out.append((k, de / ((k - 1) * four_dt)))
            k, de = 0, 0.0
    if k >= kmin:
        out.append((k, de / ((k - 1) * four_dt)))
    return out


@pytest.mark.parametrize("sim", range(6))
def test_tracks_rederived_from_trajectory(sim):
    # a thick detection slab so a 60-frame walk opens and closes many tracks
    model = small_model(Lx=4.0, Ly=4.0, Lz=4.0, ax=2.5, ay=3.0, kmin=2)
    osys, op, p = run(model, [2.0, 2.0], [1e9, 1e9], [2.4], [1], [60], sim_id=sim)
    pd = O.orc_proposal_f64()
    x32, _ = O.trace_particle(osys, op, pd, 4, 0, 0)
    want = tracks_from_trajectory(x32, 1.25, 1.5, np.float32(1.2), 2, float(np.float32(0.2)))
    K, DE = O.tracks_f32(osys, op, 4)
    assert [k for k, _ in want] == K.tolist()
    assert np.allclose([v for _, v in want], DE, rtol=2e-5, atol=1e-6)


@pytest.mark.parametrize("box", [(200.0, 200.0, 30.0, 160.0, 120.0), (230.0, 140.0, 26.0, 150.0, 70.0)])
def test_fp64_twin_and_drift(box):
    """The FP32 walk (box units) against its FP64 twin (um, libm), in the
    headline box and in one where every axis has its own length and window."""
    Lx, Ly, Lz, ax, ay = box
    model = cases.headline_model(Lx=Lx, Ly=Ly, Lz=Lz, ax=ax, ay=ay)
    p, f, n = cases.fixed_proposal(model, [1200], [1000], sim_id=11)
    osys = O.system_from_abi(model.to_abi())
    op = O.OracleProposal(p, f, n)
    pd = O.orc_proposal_f64()
    for s, D in enumerate([cases.D175, cases.D500]):
        pd.sigma[s] = math.sqrt(2 * D * 0.05)
        pd.half_az[s] = 1.0
        pd.cum_prob[s] = [0.5, 2.0][s]
    H32, t32, _ = O.simulate_f32(osys, op, 1)
    H64, t64, _ = O.simulate_f64(osys, op, pd, 20.0, 0.05, 1)
    diff = np.abs(H32.astype(np.int64) - H64).sum()
    assert diff <= max(4, 0.005 * t32), (diff, t32)
    # FP32 vs FP64 position drift over 1000 frames (modulo the box)
    worst = 0.0
    for part in range(0, 1200, 60):
        x32, x64 = O.trace_particle(osys, op, pd, 1, 0, part)
        dlt = np.abs(x32.astype(np.float64) - x64)
        L = np.array([Lx, Ly, Lz])
        dlt = np.minimum(dlt, L - dlt)
        worst = max(worst, dlt.max())
    assert worst < 5e-3, worst


def test_tracks_longer_than_rows_are_dropped():
    model = small_model(ax=200.0, ay=200.0, n_species=1, n_k=10, kmin=2)
    osys, op, p = run(model, [1.0], [1e9], [60.0], [100], [20])
    H, total, dropped = O.simulate_f32(osys, op, 1)
    assert total == 100 and dropped == 100 and H.sum() == 0


@pytest.mark.parametrize("gauss", [False, True])
def test_reflecting_and_periodic_z_have_the_same_histogram_law(gauss):
    """The reference wraps z periodically; the model can also reflect it. For
    a window centred in z (slab, or Gaussian detection about the centre) the
    two give histograms of the same law: unfolding the reflecting walk onto a
    circle of twice the box depth turns it into a free walk whose distance to
    the box centre is that of the periodic walk. Independent seeds, 120
    simulations each: track counts, mean K and mean DE bin agree within
    4 standard errors."""
    from part_conc_spt_abc_amd import _abi

    kw = dict(detect_model=_abi.PCS_DETECT_GAUSS_Z, detect_sigma=1.0) if gauss else {}
    stats_by_z = {}
    for zb in (_abi.PCS_Z_PERIODIC, _abi.PCS_Z_REFLECT):
        m = small_model(n_k=200, z_boundary=zb, **kw)
        osys = O.system_from_abi(m.to_abi())
        rows = []
        for trial in range(120):
            p = proposal_record(m, [cases.D175, cases.D500], [5e8, 5e8], [2.0], trial, 0, 1)
            op = O.OracleProposal(p, np.array([200], np.uint32), np.array([300], np.uint32))
            H, total, _ = O.simulate_f32(osys, op, 7919 * (zb + 1) + trial)
            n = max(int(H.sum()), 1)
            rows.append((total, (H.sum(axis=1) * np.arange(1, m.n_k + 1)).sum() / n,
                         (H.sum(axis=0) * np.arange(m.n_de)).sum() / n))
        stats_by_z[zb] = np.array(rows, np.float64)
    a, b = stats_by_z[_abi.PCS_Z_PERIODIC], stats_by_z[_abi.PCS_Z_REFLECT]
    se = np.sqrt(a.var(axis=0) / len(a) + b.var(axis=0) / len(b))
    z = (a.mean(axis=0) - b.mean(axis=0)) / se
    assert np.all(np.abs(z) < 4.0), z
