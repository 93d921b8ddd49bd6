"""GPU parity: the HIP path through the C ABI against the CPU oracle.

Integer outputs (histograms, track counts, exact distance sums) must be
bit-exact; the device FP32 math must equal the oracle's bit for bit.
"""
import math

import numpy as np
import pytest

import cases
import oracle_lib as O
from part_conc_spt_abc_amd import _abi
from part_conc_spt_abc_amd.gpu import Context, selftest_math, selftest_sqrt

pytestmark = pytest.mark.gpu


def oracle_hists(model, props, frames, particles, seed):
    osys = O.system_from_abi(model.to_abi())
    out = []
    for p, (f, n) in zip(props, cases.split_videos(props, frames, particles)):
        H, total, dropped = O.simulate_f32(osys, O.OracleProposal(p, f, n), seed)
        out.append((H, total - dropped))
    return out


def run_gpu(ctx, props, frames, particles, seed, mode):
    b = ctx.batch()
    b.load(props, frames, particles)
    b.run(seed, mode)
    dist, tracks, H = b.read(hist=True)
    exact = b.read_exact()
    reruns = b.overflow_reruns()
    b.close()
    return dist, tracks, H, exact, reruns


def check_against_oracle(model, obs, props, frames, particles, seed, dist, tracks, H, exact):
    ref = oracle_hists(model, props, frames, particles, seed)
    for i, (Ho, n_ok) in enumerate(ref):
        assert np.array_equal(H[i], Ho), f"proposal {i}: histogram differs ({np.abs(H[i].astype(np.int64) - Ho).sum()} counts)"
        assert tracks[i] == n_ok, f"proposal {i}: tracks {tracks[i]} != {n_ok}"
        ex, lg = O.distance(Ho, obs, model.dist_kind)
        assert int(exact[i]) == ex, f"proposal {i}: exact distance {int(exact[i])} != {ex}"
        if ex > 0:
            assert dist[i] == pytest.approx(math.log10(ex), rel=0, abs=1e-12)


def test_device_math_bit_exact_all_inputs(gpu_lib):
    # Every radius input (the top 24 bits of a word) and every angle code
    # (the top 16 bits) through the device functions of the walk.
    w = (np.arange(1 << 24, dtype=np.uint64) << 8).astype(np.uint32)
    r, s, c = selftest_math(w)
    ro = O.bm_radius(w)
    so, co = O.sincos_angle(w)
    assert np.array_equal(r.view(np.uint32), ro.view(np.uint32))
    assert np.array_equal(s.view(np.uint32), so.view(np.uint32))
    assert np.array_equal(c.view(np.uint32), co.view(np.uint32))


def test_device_sqrt_correctly_rounded_on_its_range(gpu_lib):
    """The walk's square root (rsq + one Markstein correction) is correctly
    rounded on every float in [2^-102, 64), which holds every input the log
    gives it: the smallest, -2 ln u at u = 1, is the tiny term
    2^-100 Q(0) = 0x1.00006p-102, the largest 33.3. So the oracle's sqrtf
    matches it. Far below 2^-102, where the correction's residual term is
    subnormal, it misses."""
    bits = lambda v: int(np.array([v], np.float32).view(np.uint32)[0])
    y1 = 2.0 ** -100 * (-float.fromhex("-0x1.000060p-1") / 2.0)
    assert np.float32(y1) == y1 and 2.0 ** -102 <= y1
    assert selftest_sqrt(bits(2.0 ** -102), bits(64.0)) == 0
    assert selftest_sqrt(bits(2.0 ** -126), bits(2.0 ** -118)) > 0


FIVE_CLASSES = [cases.stokes_einstein(d) for d in (100, 200, 300, 450, 600)]


def observed(model, seed=0, particles=1200, frames=None):
    """Data histogram from the oracle at the true parameters (seed 0)."""
    frames = frames or [model.n_k]
    m = FIVE_CLASSES[: model.n_species] if model.n_species != 2 else None
    p, f, n = cases.fixed_proposal(model, [particles] * len(frames), frames, m=m, sim_id=0xFFFFFFF0)
    H, _, _ = O.simulate_f32(O.system_from_abi(model.to_abi()), O.OracleProposal(p, f, n), seed)
    return H


@pytest.mark.parametrize("mode", [_abi.PCS_MODE_PLAIN, _abi.PCS_MODE_BATCHED, _abi.PCS_MODE_BATCHED_DIRECT,
                                  _abi.PCS_MODE_PLAIN_STREAMS])
def test_headline_single_simulation(gpu_lib, mode):
    model = cases.cpu_model()  # 500 frames, H 500 x 200
    obs = observed(model)
    p, f, n = cases.fixed_proposal(model, [1200], [500], sim_id=7)
    with Context(model, obs) as ctx:
        dist, tracks, H, exact, _ = run_gpu(ctx, [p], f, n, 1, mode)
    print(f"mode {mode}: tracks {tracks[0]} dist {dist[0]:.6f}", flush=True)
    check_against_oracle(model, obs, [p], f, n, 1, dist, tracks, H, exact)


@pytest.mark.parametrize("mode", [_abi.PCS_MODE_PLAIN, _abi.PCS_MODE_BATCHED])
def test_maximum_histogram_shape(gpu_lib, mode):
    """The ABI's largest histogram, n_k = n_de = 4096 (the distance's widest
    column segments and its largest LDS), with rows dense and sparse in
    records, and a proposal whose particles never leave the window (slab
    deeper than the box, full lateral window), so every track outlasts n_k
    and is dropped; against the oracle."""
    model = cases.headline_model(n_k=_abi.PCS_MAX_NK, n_de=_abi.PCS_MAX_NDE, ax=200.0, ay=200.0)
    obs = observed(model, particles=800, frames=[2000])
    m, c = [cases.D175, cases.D500], [5e8, 5e8]
    props = [cases.proposal_record(model, m, c, [2.0], 1, 0, 1),
             cases.proposal_record(model, m, c, [2.0], 2, 1, 2),
             cases.proposal_record(model, m, c, [60.0], 3, 3, 1)]
    f = np.array([4100, 1500, 300, 4100], np.uint32)
    n = np.array([2500, 3000, 500, 40], np.uint32)
    with Context(model, obs) as ctx:
        dist, tracks, H, exact, _ = run_gpu(ctx, props, f, n, 17, mode)
    assert tracks[2] == 0 and not H[2].any()
    check_against_oracle(model, obs, props, f, n, 17, dist, tracks, H, exact)


def test_batched_mixed_sizes(gpu_lib):
    model = cases.headline_model()  # H 1000 x 500
    obs = observed(model)
    rng = np.random.default_rng(3)
    props, f, n = cases.draw_proposals(model, 24, rng, frames_choices=(200, 500, 1000, 2000), c_range=(1e8, 2e9))
    print("particles", n.tolist(), "frames", f.tolist(), flush=True)
    with Context(model, obs) as ctx:
        for mode in (_abi.PCS_MODE_BATCHED, _abi.PCS_MODE_PLAIN, _abi.PCS_MODE_PLAIN_STREAMS):
            dist, tracks, H, exact, _ = run_gpu(ctx, props, f, n, 11, mode)
            check_against_oracle(model, obs, props, f, n, 11, dist, tracks, H, exact)


def test_full_generation_properties(gpu_lib):
    """One generation at its full size: 100,000 proposals, total c log-uniform
    on 1e8..1e10 /mL (120-12,000 particles), frames 500/1,000/2,000, H 1000 x
    500. Graph replay, direct launches and a repeat agree bit for bit; every
    distance is log10 of its exact sum; proposals taken out of the generation
    and simulated alone in plain launches reproduce their tracks and exact
    distances, and their histograms equal the oracle's."""
    model = cases.headline_model()
    obs = observed(model)
    rng = np.random.default_rng(2024)
    props, f, n = cases.draw_proposals(model, 100_000, rng, frames_choices=(500, 1000, 2000))
    steps = n.astype(np.int64) * f
    pick = [int(np.argmin(steps)), int(np.argmax(steps))] + [int(i) for i in rng.choice(len(props), 6, replace=False)]
    print("sample", pick, "particles", n[pick].tolist(), "frames", f[pick].tolist(), flush=True)
    with Context(model, obs) as ctx:
        b = ctx.batch()
        b.load(props, f, n)
        runs = []
        for mode in (_abi.PCS_MODE_BATCHED, _abi.PCS_MODE_BATCHED_DIRECT, _abi.PCS_MODE_BATCHED):
            b.run(31, mode)
            dist, tracks, _ = b.read()
            runs.append((dist, tracks, b.read_exact()))
        b.close()
        dist, tracks, exact = runs[0]
        for d2, t2, e2 in runs[1:]:
            assert np.array_equal(exact, e2) and np.array_equal(tracks, t2) and np.array_equal(dist, d2)
        pos = exact > 0
        assert pos.mean() > 0.99
        np.testing.assert_allclose(dist[pos], np.log10(exact[pos].astype(np.float64)), rtol=0, atol=1e-12)
        assert tracks[int(np.argmax(steps))] > tracks[int(np.argmin(steps))]

        sub = []
        for j, i in enumerate(pick):
            q = _abi.pcs_proposal.from_buffer_copy(props[i])
            q.first_video = j
            sub.append(q)
        fs, ns = f[pick], n[pick]
        sd, st, sH, se, _ = run_gpu(ctx, sub, fs, ns, 31, _abi.PCS_MODE_PLAIN)
    assert np.array_equal(se, exact[pick]) and np.array_equal(st, tracks[pick])
    check_against_oracle(model, obs, sub, fs, ns, 31, sd, st, sH, se)


@pytest.mark.parametrize("tune", [dict(particles_per_lane=2), dict(waves_per_group=1, stage_records=128),
                                  dict(waves_per_group=2, groups_per_cu=2, stage_records=512),
                                  dict(particles_per_lane=2, gauss=True)])
def test_tuning_does_not_change_results(gpu_lib, tune):
    tune = dict(tune)
    kw = dict(detect_model=_abi.PCS_DETECT_GAUSS_Z, detect_sigma=1.0) if tune.pop("gauss", False) else {}
    model = cases.headline_model(n_k=600, n_de=300, **kw)
    obs = observed(model, frames=[600])
    rng = np.random.default_rng(5)
    props, f, n = cases.draw_proposals(model, 10, rng, frames_choices=(300, 600), videos=3, c_range=(1e8, 1e9))
    with Context(model, obs) as ctx:
        ctx.set_tuning(**tune)
        dist, tracks, H, exact, _ = run_gpu(ctx, props, f, n, 21, _abi.PCS_MODE_BATCHED)
    check_against_oracle(model, obs, props, f, n, 21, dist, tracks, H, exact)


@pytest.mark.parametrize("variant", ["reflect", "gauss", "five_species_l1", "variable_depth", "non_square",
                                     "non_square_gauss_reflect"])
def test_model_variants(gpu_lib, variant):
    kw = {}
    m = az = None
    if variant == "reflect":
        kw = dict(z_boundary=_abi.PCS_Z_REFLECT)
    elif variant == "gauss":
        kw = dict(detect_model=_abi.PCS_DETECT_GAUSS_Z, detect_sigma=1.0)
    elif variant.startswith("non_square"):
        # every axis its own length and window, so a swapped axis shows
        kw = dict(Lx=230.0, Ly=140.0, Lz=26.0, ax=150.0, ay=70.0)
        if variant.endswith("gauss_reflect"):
            kw.update(detect_model=_abi.PCS_DETECT_GAUSS_Z, detect_sigma=1.5, z_boundary=_abi.PCS_Z_REFLECT)
    elif variant == "five_species_l1":
        kw = dict(n_species=5, dist_kind=_abi.PCS_DIST_L1)
        m = FIVE_CLASSES
    model = cases.headline_model(n_k=800, n_de=400, **kw)
    obs = observed(model, frames=[800])
    props = []
    fr, pa = [], []
    rng = np.random.default_rng(9)
    for i in range(6):
        C = model.n_species
        mm = m or [cases.D175, cases.D500]
        c = list(rng.dirichlet(np.ones(C)) * 1e9)
        az = [float(x) for x in rng.uniform(0.5, 5.0, C)] if variant == "variable_depth" else [2.0]
        props.append(cases.proposal_record(model, mm, c, az, 100 + i, len(fr), 2))
        fr += [800, 400]
        pa += [int(x) for x in rng.integers(50, 2000, 2)]
    fr = np.array(fr, np.uint32)
    pa = np.array(pa, np.uint32)
    with Context(model, obs) as ctx:
        dist, tracks, H, exact, _ = run_gpu(ctx, props, fr, pa, 5, _abi.PCS_MODE_BATCHED)
    check_against_oracle(model, obs, props, fr, pa, 5, dist, tracks, H, exact)


def test_overflow_reruns_are_exact(gpu_lib):
    model = cases.cpu_model()
    obs = observed(model)
    rng = np.random.default_rng(1)
    props, f, n = cases.draw_proposals(model, 8, rng, frames_choices=(500,), c_range=(5e8, 2e9))
    with Context(model, obs) as ctx:
        ctx.set_record_capacity_factor(0.01)
        dist, tracks, H, exact, reruns = run_gpu(ctx, props, f, n, 2, _abi.PCS_MODE_BATCHED)
    assert reruns > 0
    check_against_oracle(model, obs, props, f, n, 2, dist, tracks, H, exact)


def test_graph_replays_with_eager_work_between(gpu_lib):
    model = cases.cpu_model()
    obs = observed(model)
    rng = np.random.default_rng(4)
    props, f, n = cases.draw_proposals(model, 6, rng, frames_choices=(500,), c_range=(1e8, 1e9))
    with Context(model, obs) as ctx:
        b = ctx.batch()
        b.load(props, f, n)
        for seed in (3, 4, 3):
            b.run(seed, _abi.PCS_MODE_BATCHED)
            dist, tracks, H = b.read(hist=True)
            exact = b.read_exact()
            # eager work on the same context between replays
            dist2, _, _, _, _ = run_gpu(ctx, props[:2], f[:2], n[:2], seed + 100, _abi.PCS_MODE_PLAIN)
            check_against_oracle(model, obs, props, f, n, seed, dist, tracks, H, exact)
        b.close()


@pytest.mark.parametrize("kmin", [2, 7, 40])
def test_kmin_values(gpu_lib, kmin):
    """Shortest recorded track from the smallest allowed (2: one step, one
    squared displacement) to one far above the typical track, over videos
    whose step counts fill the last 4-step chunk (96) or leave it partial (499)."""
    model = cases.headline_model(n_k=500, n_de=200, kmin=kmin)
    obs = observed(model, frames=[500])
    rng = np.random.default_rng(40 + kmin)
    props, f, n = cases.draw_proposals(model, 6, rng, frames_choices=(97, 500), c_range=(1e8, 1e9))
    with Context(model, obs) as ctx:
        for mode in (_abi.PCS_MODE_PLAIN, _abi.PCS_MODE_BATCHED):
            dist, tracks, H, exact, _ = run_gpu(ctx, props, f, n, 31, mode)
            check_against_oracle(model, obs, props, f, n, 31, dist, tracks, H, exact)


def test_empty_and_tiny_inputs(gpu_lib):
    model = cases.cpu_model()
    obs = observed(model)
    # zero particles, one frame, kmin-length videos
    props, fr, pa = [], [], []
    for i, (np_, nf) in enumerate([(0, 500), (5, 1), (64, 3), (65, 2), (1, 500)]):
        p, f, n = cases.fixed_proposal(model, [np_], [nf], sim_id=i)
        p.first_video = len(fr)
        props.append(p)
        fr += list(f)
        pa += list(n)
    fr = np.array(fr, np.uint32)
    pa = np.array(pa, np.uint32)
    with Context(model, obs) as ctx:
        for mode in (_abi.PCS_MODE_PLAIN, _abi.PCS_MODE_BATCHED):
            dist, tracks, H, exact, _ = run_gpu(ctx, props, fr, pa, 8, mode)
            check_against_oracle(model, obs, props, fr, pa, 8, dist, tracks, H, exact)


def test_invalid_arguments_fail_loudly(gpu_lib):
    model = cases.cpu_model()
    bad = model.to_abi()
    bad.kmin = 1
    import ctypes as C

    h = C.c_void_p()
    assert gpu_lib.pcs_open(C.byref(bad), None, 0, C.byref(h)) == _abi.PCS_EINVAL
    with Context(model) as ctx:
        p, f, n = cases.fixed_proposal(model, [10], [10])
        p.cum_prob[1] = 0.5  # last cumulative weight < 1
        with pytest.raises(_abi.PcsError):
            b = ctx.batch()
            b.load([p], f, n)
        # more than 2^31 particles in a video (tile indices are 32-bit)
        q, f, _ = cases.fixed_proposal(model, [10], [10])
        with pytest.raises(_abi.PcsError) as e:
            ctx.batch().load([q], f, np.array([(1 << 31) + 1], np.uint32))
        assert e.value.code == _abi.PCS_EINVAL
    from part_conc_spt_abc_amd.gpu import Server

    srv = Server(device=0, arena_bytes=1 << 24, max_inflight=8)
    sid = srv.add_system(model)
    q, f, _ = cases.fixed_proposal(model, [10], [10])
    with pytest.raises(_abi.PcsError) as e:
        srv.submit(0, sid, q, f, np.array([(1 << 31) + 1], np.uint32), 1)
    assert e.value.code == _abi.PCS_EINVAL
    srv.close()


def test_failed_load_leaves_nothing_to_run(gpu_lib):
    """A batch whose last load failed part-way (here its second proposal is
    invalid, after the batch held a smaller valid load and ran it) refuses to
    run or read (PCS_ESTATE, on the host: no kernel is launched), and a
    following valid load runs and reads as a fresh batch would."""
    model = cases.cpu_model()
    obs = observed(model)
    good, f1, n1 = cases.fixed_proposal(model, [300], [200], sim_id=1)
    with Context(model, obs) as ctx:
        b = ctx.batch()
        b.load([good], f1, n1)
        b.run(5, _abi.PCS_MODE_BATCHED)
        d_first, t_first, _ = b.read()
        bad = _abi.pcs_proposal.from_buffer_copy(good)
        bad.first_video = 1
        bad.cum_prob[1] = 0.5
        frames = np.array([200, 200, 200], np.uint32)
        parts = np.array([300, 300, 300], np.uint32)
        with pytest.raises(_abi.PcsError):
            b.load([good, bad, good], frames, parts)
        for mode in (_abi.PCS_MODE_BATCHED, _abi.PCS_MODE_PLAIN):
            with pytest.raises(_abi.PcsError) as e:
                b.run(5, mode)
            assert e.value.code == _abi.PCS_ESTATE
        with pytest.raises(_abi.PcsError) as e:
            b.read()
        assert e.value.code == _abi.PCS_ESTATE
        b.load([good], f1, n1)
        b.run(5, _abi.PCS_MODE_BATCHED)
        d_again, t_again, _ = b.read()
        b.close()
    assert np.array_equal(t_again, t_first) and np.array_equal(d_again.view(np.uint64), d_first.view(np.uint64))


def test_batch_outlives_its_context(gpu_lib):
    # pcs_close releases the caller's handle only: a live batch keeps the
    # context, runs after the close, and destroying it frees both.
    model = cases.headline_model(n_k=300, n_de=200)
    obs = observed(model, frames=[300])
    rng = np.random.default_rng(13)
    props, f, n = cases.draw_proposals(model, 6, rng, frames_choices=(300,), c_range=(1e8, 1e9))
    ctx = Context(model, obs)
    b = ctx.batch()
    b.load(props, f, n)
    ctx.close()
    b.run(41, _abi.PCS_MODE_BATCHED)
    dist, tracks, H = b.read(hist=True)
    exact = b.read_exact()
    b.close()
    check_against_oracle(model, obs, props, f, n, 41, dist, tracks, H, exact)
    # and later launches on this thread report no stale error
    with Context(model, obs) as ctx2:
        dist2, tracks2, _, _, _ = run_gpu(ctx2, props, f, n, 41, _abi.PCS_MODE_BATCHED)
    assert np.array_equal(dist2, dist) and np.array_equal(tracks2, tracks)


def test_autotune_picks_a_valid_shape_and_keeps_results(gpu_lib):
    model = cases.headline_model(n_k=600, n_de=300)
    obs = observed(model, frames=[600])
    rng = np.random.default_rng(12)
    props, f, n = cases.draw_proposals(model, 40, rng, frames_choices=(300, 600), c_range=(1e8, 2e9))
    with Context(model, obs) as ctx:
        best, ms = ctx.autotune(props, f, n)
        assert best["particles_per_lane"] in (1, 2) and best["waves_per_group"] in (2, 4)
        assert best["waves_per_group"] * best["groups_per_cu"] <= 32 and ms > 0
        assert ctx.get_tuning() == best
        dist, tracks, H, exact, _ = run_gpu(ctx, props, f, n, 33, _abi.PCS_MODE_BATCHED)
    check_against_oracle(model, obs, props, f, n, 33, dist, tracks, H, exact)


def test_large_polydisperse_model(gpu_lib):
    """The large polydisperse configuration (400x400x60 um box, 5 classes
    100-600 nm with Dirichlet abundances, Gaussian z detection sigma 1 um,
    2000 frames, H 2000 x 1000) with 12,000 particles so the oracle finishes
    in seconds; FP32 drift against the FP64 twin is bounded."""
    model = cases.headline_model(n_k=2000, n_de=1000, Lx=400.0, Ly=400.0, Lz=60.0, n_species=5,
                                 detect_model=_abi.PCS_DETECT_GAUSS_Z, detect_sigma=1.0)
    rng = np.random.default_rng(50)
    w = rng.dirichlet(np.ones(5))
    p = cases.proposal_record(model, FIVE_CLASSES, list(w * 1e10), [2.0], 77, 0, 1)
    f, n = np.array([2000], np.uint32), np.array([12000], np.uint32)
    obs = observed(model, particles=3000, frames=[2000])
    with Context(model, obs) as ctx:
        dist, tracks, H, exact, _ = run_gpu(ctx, [p], f, n, 6, _abi.PCS_MODE_BATCHED)
    check_against_oracle(model, obs, [p], f, n, 6, dist, tracks, H, exact)
    osys = O.system_from_abi(model.to_abi())
    op = O.OracleProposal(p, f, n)
    pd = O.orc_proposal_f64()
    acc = 0.0
    for i, D in enumerate(FIVE_CLASSES):
        pd.sigma[i] = math.sqrt(2 * D * 0.05)
        pd.half_az[i] = 1.0
        acc += w[i]
        pd.cum_prob[i] = acc if i < 4 else 2.0
    worst = 0.0
    for part in range(0, 12000, 1500):
        x32, x64 = O.trace_particle(osys, op, pd, 6, 0, part)
        d = np.abs(x32.astype(np.float64) - x64)
        L = np.array([400.0, 400.0, 60.0])
        worst = max(worst, np.minimum(d, L - d).max())
    print(f"FP32 vs FP64 position drift over 2000 frames: {worst:.2e} um", flush=True)
    assert worst < 1e-2


def test_large_polydisperse_model_full_size(gpu_lib):
    """The same configuration at its full size: c = 1e10 /mL in the 400x400x60
    um box, a Poisson(96,000) particle count (95,727 particles x 2000 frames),
    bit-equal to the oracle in the batched and plain forms (the oracle takes
    10-20 s)."""
    model = cases.headline_model(n_k=2000, n_de=1000, Lx=400.0, Ly=400.0, Lz=60.0, n_species=5,
                                 detect_model=_abi.PCS_DETECT_GAUSS_Z, detect_sigma=1.0)
    rng = np.random.default_rng(51)
    w = rng.dirichlet(np.ones(5))
    n = int(rng.poisson(1e10 * model.volume_factor))
    assert n == 95727
    p = cases.proposal_record(model, FIVE_CLASSES, list(w * 1e10), [2.0], 78, 0, 1)
    f, n = np.array([2000], np.uint32), np.array([n], np.uint32)
    obs = observed(model, particles=3000, frames=[2000])
    with Context(model, obs) as ctx:
        runs = [run_gpu(ctx, [p], f, n, 8, mode) for mode in (_abi.PCS_MODE_BATCHED, _abi.PCS_MODE_PLAIN)]
    for dist, tracks, H, exact, _ in runs[1:]:
        assert np.array_equal(H, runs[0][2]) and np.array_equal(exact, runs[0][3])
    dist, tracks, H, exact, _ = runs[0]
    print("tracks", int(tracks[0]), flush=True)
    check_against_oracle(model, obs, [p], f, n, 8, dist, tracks, H, exact)


def test_generate_tracks_matches_oracle(gpu_lib):
    """Track-list form (generate_experiment.jl) on the GPU: the same (K, DE)
    multiset as the oracle, DE bit for bit."""
    import ctypes as C

    for variant in ("slab", "gauss", "kmin2"):
        kw = {"slab": {}, "gauss": dict(detect_model=_abi.PCS_DETECT_GAUSS_Z, detect_sigma=1.0),
              "kmin2": dict(kmin=2)}[variant]  # kmin 2: tracks of one step, divisor (k - 1) 4 dt = 4 dt
        model = cases.headline_model(**kw)
        p, f, n = cases.fixed_proposal(model, [900, 0, 130], [1000, 500, 2000], sim_id=41)
        with Context(model) as ctx:
            cap = 1 << 20
            K = np.zeros(cap, np.uint32)
            DE = np.zeros(cap, np.float32)
            cnt = C.c_uint64(0)
            _abi.check(gpu_lib.pcs_generate_tracks(ctx.h, C.byref(p), f.ctypes.data, n.ctypes.data, len(f), 19,
                                                   K.ctypes.data, DE.ctypes.data, cap, C.byref(cnt)),
                       "pcs_generate_tracks")
        m = int(cnt.value)
        Ko, DEo = O.tracks_f32(O.system_from_abi(model.to_abi()), O.OracleProposal(p, f, n), 19)
        assert m == len(Ko) > 100
        got = sorted(zip(K[:m].tolist(), DE[:m].view(np.uint32).tolist()))
        want = sorted(zip(Ko.tolist(), DEo.view(np.uint32).tolist()))
        assert got == want, variant


@pytest.mark.parametrize("P", [1, 4, 5, 8, 9, 15])
def test_pmc_weights_on_device(gpu_lib, P):
    """Device pmc-standard weights vs the formula in FP64 numpy
    (normpdf products with their constants, then normalisation), for
    parameter counts with their own kernel (1..8) and the general one (9,
    15), with old members from two chunks, and one new member twenty
    standard deviations out on its first row (kernel terms near e^-100)."""
    import ctypes as C

    rng = np.random.default_rng(8 + P)
    n_old, n_new = 4500, 500
    old = rng.normal(0.0, 1.0, (P, n_old)) * rng.choice([0.3, 0.5, 1.0, 2.0, 3.0], (P, 1))
    new = old[:, rng.integers(0, n_old, n_new)] + rng.normal(0.0, 0.2, (P, n_new))
    new[0, 7] += 20.0 * old[0].std()
    w_old = rng.random(n_old)
    w_old /= w_old.sum()
    tau = np.sqrt(2.0 * old.var(axis=1))
    out = np.zeros(n_new)
    _abi.check(gpu_lib.pcs_pmc_weights(0, new.ctypes.data, n_new, old.ctypes.data, w_old.ctypes.data, n_old,
                                       tau.ctypes.data, P, out.ctypes.data), "pcs_pmc_weights")
    d = (new[:, :, None] - old[:, None, :]) / tau[:, None, None]
    k = np.prod(np.exp(-0.5 * d ** 2) / (np.sqrt(2 * np.pi) * tau[:, None, None]), axis=0)
    ws = 1.0 / (k * w_old[None, :]).sum(axis=1)
    ws /= ws.sum()
    assert np.allclose(out, ws, rtol=1e-11, atol=0)
    assert abs(out.sum() - 1.0) < 1e-12


@pytest.mark.parametrize("mode", [_abi.PCS_MODE_BATCHED, _abi.PCS_MODE_PLAIN])
def test_distance_sum_past_two_to_the_64(gpu_lib, mode):
    """Observed counts near 2^32 in every cumulative cell: the sum of squares
    passes 2^64. The exact value saturates at 2^64 - 1 and the log10 distance
    is the oracle's, taken from the full 128-bit sum (checked against Python
    integers)."""
    model = cases.cpu_model()
    obs = np.zeros((model.n_k, model.n_de), np.uint32)
    obs[0, 0] = 4_000_000_000
    obs[5, 7] = 150_000_000
    p, f, n = cases.fixed_proposal(model, [600], [300], sim_id=3)
    with Context(model, obs) as ctx:
        dist, tracks, H, exact, _ = run_gpu(ctx, [p], f, n, 1, mode)
    Ho, _, _ = O.simulate_f32(O.system_from_abi(model.to_abi()), O.OracleProposal(p, f, n), 1)
    assert np.array_equal(H[0], Ho)
    ex, lg = O.distance(Ho, obs, 0)
    S = np.cumsum(np.cumsum(Ho.astype(object), axis=0), axis=1)
    T = np.cumsum(np.cumsum(obs.astype(object), axis=0), axis=1)
    want = sum(int(v) * int(v) for v in (S - T).ravel())
    assert want > 2 ** 64 and ex == 2 ** 64 - 1 and int(exact[0]) == ex
    assert dist[0] == pytest.approx(lg, rel=1e-15) and lg == pytest.approx(math.log10(want), rel=1e-15)
