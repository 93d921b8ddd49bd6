"""Serving form against the CPU oracle, at the serving workload's full size.

The serving workload: 32 concurrent chains, each with its own box (lateral
100-400 um, depth 10-50 um, window 0.8 x 0.6 of the box) and 2-4 species of
100-600 nm, H 1000 x 500 uint32, 1,000 frames, total concentration
log-uniform on 1e8..1e10 /mL with at most 60,000 particles per request. Here
the chains also cycle through the three model classes (periodic slab,
reflecting z, Gaussian z detection). Every serving form (host-driven launch
groups; the device-resident loop, whose small requests' tiles are each walked
by a whole group, and the same loop with every tile walked by one wave) must
give every request the same (exact distance sum, tracks) alone, mixed with the
others and as a resident batch, and a sample of requests of every class must equal the oracle's
orc_simulate_f32 + orc_distance. Record slices forced to overflow are rerun
and still equal the oracle.
"""
import concurrent.futures as cf
import math
import time

import numpy as np
import pytest

import oracle_lib as O
from part_conc_spt_abc_amd import _abi
from part_conc_spt_abc_amd.gpu import Context, Server
from part_conc_spt_abc_amd.model import Model, proposal_record, stokes_einstein

pytestmark = pytest.mark.gpu

CLASSES = ("periodic slab", "reflecting z", "gaussian z")


def serving_chains(rng, n=32, n_k=1000, n_de=500):
    out = []
    for c in range(n):
        L = float(rng.uniform(100, 400))
        Lz = float(rng.uniform(10, 50))
        C = int(rng.integers(2, 5))
        kw = {}
        if c % 3 == 1:
            kw = dict(z_boundary=_abi.PCS_Z_REFLECT)
        elif c % 3 == 2:
            kw = dict(detect_model=_abi.PCS_DETECT_GAUSS_Z, detect_sigma=1.0)
        m = Model(Lx=L, Ly=L, Lz=Lz, ax=0.8 * L, ay=0.6 * L, deltat=0.05, kmin=3, n_k=n_k, n_de=n_de, ub_de=20.0,
                  n_species=C, **kw)
        D = sorted(stokes_einstein(float(d)) for d in rng.uniform(100, 600, C))
        # observed histogram: short tracks dense, a sparse tail of long ones
        obs = np.zeros((m.n_k, m.n_de), np.uint32)
        obs[2:60, :120] = rng.integers(0, 6, (58, 120))
        rows = rng.integers(60, m.n_k, 40)
        obs[rows, rng.integers(0, m.n_de, 40)] += 1
        out.append((m, D, obs))
    return out


def serving_requests(rng, ch, per_chain, frames=1000, max_particles=60000):
    reqs = []
    for i in range(per_chain * len(ch)):
        ci = i % len(ch)
        m, D, _ = ch[ci]
        ct = 10.0 ** rng.uniform(8.0, 10.0)
        w = rng.dirichlet(np.ones(m.n_species))
        p = proposal_record(m, D, list(w * ct), [2.0], i, 0, 1)
        n = min(int(rng.poisson(ct * m.volume_factor)), max_particles)
        reqs.append((ci, p, np.array([frames], np.uint32), np.array([n], np.uint32), 5000 + i))
    return reqs


def collect(srv, n, timeout):
    got = {}
    t0 = time.time()
    while len(got) < n:
        for c in srv.poll(1024, wait_us=2000):
            assert c.status == 0, (c.tag, c.status)
            got[c.tag] = (int(c.exact), int(c.tracks), float(c.dist))
        assert time.time() - t0 < timeout, f"only {len(got)} of {n} completions"
    return got


FORMS = ("host", "loop", "loop_one_wave")


def server(form, **kw):
    return Server(device=0, device_loop=form != "host", one_wave_tiles=form == "loop_one_wave", **kw)


def serve(ch, reqs, form, alone, cap_factor=None, arena=4 << 30):
    srv = server(form, arena_bytes=arena, max_inflight=1024)
    if cap_factor is not None:
        srv.set_record_capacity_factor(cap_factor)
    sids = [srv.add_system(m, obs) for m, _, obs in ch]
    got = {}
    if alone:
        for tag, (ci, p, f, n, seed) in enumerate(reqs):
            assert srv.submit(tag, sids[ci], p, f, n, seed)
            got.update(collect(srv, 1, 30.0))
    else:
        for tag, (ci, p, f, n, seed) in enumerate(reqs):
            while not srv.submit(tag, sids[ci], p, f, n, seed):
                srv.poll(0)
            if tag % 11 == 0:
                time.sleep(0.0003)  # later requests arrive while earlier ones run
        got = collect(srv, len(reqs), 120.0)
    reruns = srv.overflow_reruns()
    walked = srv.group_walked()
    if form != "loop":
        assert walked == 0
    print(f"  {walked} of {len(reqs)} requests group-walked, {reruns} reruns", flush=True)
    assert srv.pending() == 0
    srv.close()
    return got, reruns, walked


def oracle_result(ch, req):
    ci, p, f, n, seed = req
    m, _, obs = ch[ci]
    osys = O.system_from_abi(m.to_abi())
    H, total, dropped = O.simulate_f32(osys, O.OracleProposal(p, f, n), seed)
    ex, _ = O.distance(H, obs, m.dist_kind)
    return ex, total - dropped


def oracle_sample(ch, reqs, rng, per_class=3):
    """Per model class: the request with the most particles, the one with the
    fewest and random others."""
    pick = []
    for k in range(3):
        idx = [i for i, r in enumerate(reqs) if r[0] % 3 == k]
        by_n = sorted(idx, key=lambda i: int(reqs[i][3][0]))
        chosen = {by_n[0], by_n[-1]}
        rest = [i for i in idx if i not in chosen]
        chosen.update(int(x) for x in rng.choice(rest, per_class - 2, replace=False))
        pick.extend(sorted(chosen))
    with cf.ThreadPoolExecutor(8) as ex:  # the oracle releases the GIL inside its C calls
        res = list(ex.map(lambda i: oracle_result(ch, reqs[i]), pick))
    return dict(zip(pick, res))


def test_serving_workload_full_size_matches_oracle(gpu_lib):
    rng = np.random.default_rng(7)
    ch = serving_chains(rng)
    reqs = serving_requests(rng, ch, per_chain=6)
    # and per model class one request at the 60,000-particle maximum, on the
    # class's largest box at c = 1e10 /mL
    for k in range(3):
        ci = max(range(k, len(ch), 3), key=lambda c: ch[c][0].volume_factor)
        m, D, _ = ch[ci]
        p = proposal_record(m, D, list(np.full(m.n_species, 1e10 / m.n_species)), [2.0], len(reqs), 0, 1)
        reqs.append((ci, p, np.array([1000], np.uint32), np.array([60000], np.uint32), 5000 + len(reqs)))
    npart = np.array([int(r[3][0]) for r in reqs])
    print(f"{len(reqs)} requests, particles {npart.min()}..{npart.max()}, "
          f"{int((npart == 60000).sum())} at the 60,000 cap", flush=True)
    assert int((npart == 60000).sum()) >= 3
    results = {}
    for form in FORMS:
        for alone in (True, False):
            print(f"form={form} alone={alone}", flush=True)
            results[(form, alone)], _, walked = serve(ch, reqs, form, alone)
            if form == "loop" and alone:  # every request of at most 256 tiles (16,384 particles)
                assert walked == sum(int(r[3][0]) <= 256 * 64 for r in reqs)
    ref = results[("host", True)]
    for key, got in results.items():
        assert got == ref, key
    # each chain's requests as one resident batch of its model
    for ci, (m, _, obs) in enumerate(ch):
        mine = [(tag, r) for tag, r in enumerate(reqs) if r[0] == ci]
        with Context(m, obs) as ctx:
            for tag, (_, p, f, n, seed) in mine:
                b = ctx.batch()
                b.load([p], f, n)
                b.run(seed, _abi.PCS_MODE_BATCHED)
                dist, tracks, _ = b.read()
                exact = b.read_exact()
                b.close()
                assert (int(exact[0]), int(tracks[0]), float(dist[0])) == ref[tag], (ci, tag)
    # a sample of every model class against the oracle
    sample = oracle_sample(ch, reqs, np.random.default_rng(1))
    for i, (ex, tr) in sample.items():
        cls = CLASSES[reqs[i][0] % 3]
        print(f"request {i} ({cls}, {int(reqs[i][3][0])} particles): tracks {tr}, exact {ex}", flush=True)
        assert ref[i][:2] == (ex, tr), (i, cls)
        assert ref[i][2] == pytest.approx(math.log10(ex), rel=0, abs=1e-12)


@pytest.mark.parametrize("form", FORMS)
def test_serving_overflow_reruns_match_oracle(gpu_lib, form):
    """Record slices sized far below the tracks (capacity factor 1e-4) all
    overflow; the server reruns each request with an exact slice before it
    completes, and the results equal the oracle's."""
    rng = np.random.default_rng(13)
    ch = serving_chains(rng, n=3, n_k=500, n_de=200)
    reqs = serving_requests(rng, ch, per_chain=3, frames=400, max_particles=6000)
    got, reruns, _ = serve(ch, reqs, form, alone=False, cap_factor=1e-4, arena=1 << 30)
    print(f"reruns {reruns}", flush=True)
    assert reruns > 0
    with cf.ThreadPoolExecutor(8) as ex:
        ref = list(ex.map(lambda r: oracle_result(ch, r), reqs))
    for i, (e, t) in enumerate(ref):
        assert got[i][:2] == (e, t), i
