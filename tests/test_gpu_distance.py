"""The distance kernel over every size of record slice, against the oracle's
distance (orc_distance) of the same histogram.

A workgroup reads a slice of T records once when T <= 8,192 (held in
registers), sorts the 16-bit bins by row into LDS when they fit there, and
otherwise into the slice's HBM scratch; past 8,192 records the rest of the
slice is read a second time. Record counts from 0 to about 40,000 per
simulation, with a cluster just below 8,192, cross every one of those boundaries at
H 500x200, 1000x500 and 1000x2000 (three column-segment widths and three LDS
capacities). The batched and plain forms, and both serving forms, return the
exact sum of the histogram the batch exports."""
import math
import time

import numpy as np
import pytest

import cases
import oracle_lib as O
from part_conc_spt_abc_amd import _abi
from part_conc_spt_abc_amd.gpu import Context, Server

pytestmark = pytest.mark.gpu

FRAMES = 1000
HELD = 8192  # records a distance workgroup holds in registers


def observed(model):
    p, f, n = cases.fixed_proposal(model, [1500], [FRAMES], sim_id=0xFFFFFFF0)
    H, _, _ = O.simulate_f32(O.system_from_abi(model.to_abi()), O.OracleProposal(p, f, n), 0)
    return H


# Record-count bands: LDS-sorted; held in registers with the bins sorted in
# LDS or (past the LDS capacity of the launch, about 7,200 at H 1000x500) in
# HBM; past the registers (read twice), also with bins in HBM.
BANDS = [(1, 4096), (7300, 8192), (8193, 16384), (16385, 10 ** 9)]


def sweep(model, n_sims=64, most=14000):
    """One video of FRAMES frames per simulation: particles 0..most, plus a
    cluster at 2,400..2,800 particles (about 2.98 records per particle over
    1,000 frames puts it across 7,300..8,192 records)."""
    parts = np.concatenate([np.linspace(0, most, n_sims), np.linspace(2400, 2800, 9)]).astype(np.uint32)
    n_sims = len(parts)
    props = []
    for i in range(n_sims):
        p, _, _ = cases.fixed_proposal(model, [int(parts[i])], [FRAMES], sim_id=100 + i)
        p.first_video = i
        props.append(p)
    return props, np.full(n_sims, FRAMES, np.uint32), parts


@pytest.mark.parametrize("n_k,n_de", [(500, 200), (1000, 500), (1000, 2000)])
def test_distance_over_record_counts(gpu_lib, n_k, n_de):
    model = cases.headline_model(n_k=n_k, n_de=n_de)
    obs = observed(model)
    props, f, n = sweep(model)
    with Context(model, obs) as ctx:
        out = {}
        for mode in (_abi.PCS_MODE_BATCHED, _abi.PCS_MODE_PLAIN):
            b = ctx.batch()
            b.load(props, f, n)
            b.run(41, mode)
            dist, tracks, H = b.read(hist=True)
            out[mode] = (dist.copy(), tracks.copy(), H.copy(), b.read_exact().copy())
            b.close()
    dist, tracks, H, exact = out[_abi.PCS_MODE_BATCHED]
    print(f"H {n_k}x{n_de}: records per simulation {tracks.min()}..{tracks.max()}", flush=True)
    assert tracks.min() == 0 and tracks.max() > 4 * HELD
    for lo, hi in BANDS:
        assert ((tracks >= lo) & (tracks <= hi)).any(), (lo, hi)
    d2, t2, H2, e2 = out[_abi.PCS_MODE_PLAIN]
    assert np.array_equal(t2, tracks) and np.array_equal(e2, exact) and np.array_equal(H2, H)
    for i in range(len(props)):
        assert int(H[i].sum()) == int(tracks[i])
        ex, _ = O.distance(H[i], obs, model.dist_kind)
        assert int(exact[i]) == ex, f"simulation {i} ({tracks[i]} records): {int(exact[i])} != {ex}"
        assert dist[i] == pytest.approx(math.log10(ex), rel=0, abs=1e-12)


@pytest.mark.parametrize("device_loop,one_wave", [(False, False), (True, False), (True, True)])
def test_serving_distance_over_record_counts(gpu_lib, device_loop, one_wave):
    """The serving forms on the same sweep at H 1000x500 (the request loop
    sorts into LDS up to its own capacity and re-reads the records; its
    tiles walked by whole groups or by one wave each)."""
    model = cases.headline_model()
    obs = observed(model)
    props, f, n = sweep(model, n_sims=40)
    with Context(model, obs) as ctx:
        b = ctx.batch()
        b.load(props, f, n)
        b.run(43, _abi.PCS_MODE_BATCHED)
        dist, tracks, _ = b.read()
        exact = b.read_exact()
        b.close()
    srv = Server(device=0, arena_bytes=2 << 30, max_inflight=64, device_loop=device_loop, one_wave_tiles=one_wave)
    sid = srv.add_system(model, obs)
    for i, p in enumerate(props):
        q = _abi.pcs_proposal.from_buffer_copy(p)
        q.first_video = 0
        assert srv.submit(i, sid, q, f[i:i + 1], n[i:i + 1], 43)
    got = {}
    t0 = time.time()
    while len(got) < len(props):
        for c in srv.poll(256, wait_us=2000):
            assert c.status == 0
            got[c.tag] = (float(c.dist), int(c.tracks), int(c.exact))
        assert time.time() - t0 < 120.0, f"only {len(got)} of {len(props)} completions"
    srv.close()
    for i in range(len(props)):
        assert got[i] == (float(dist[i]), int(tracks[i]), int(exact[i])), i


@pytest.mark.parametrize("obs_kind", ["empty", "row3"])
def test_distance_past_the_deepest_observed_row(gpu_lib, obs_kind):
    """Past the deepest observed row the sweep visits only the rows holding
    records (found 64 rows at a time) and adds each run of empty rows as a
    multiple of the row above. An empty observed histogram (every row is past
    it) and one with tracks only in row 3 put every kind of row there: dense
    rows (more than 32 records) of the large simulations, sparse rows, and
    gaps of every length; batched and in the request loop."""
    model = cases.headline_model()
    obs = np.zeros((model.n_k, model.n_de), np.uint32)
    if obs_kind == "row3":
        obs[3, :40] = np.arange(1, 41, dtype=np.uint32)
    props, f, n = sweep(model, n_sims=32)
    with Context(model, obs) as ctx:
        b = ctx.batch()
        b.load(props, f, n)
        b.run(47, _abi.PCS_MODE_BATCHED)
        dist, tracks, H = b.read(hist=True)
        exact = b.read_exact().copy()
        b.close()
    dense = 0
    for i in range(len(props)):
        ex, _ = O.distance(H[i], obs, model.dist_kind)
        assert int(exact[i]) == ex, f"simulation {i} ({tracks[i]} records): {int(exact[i])} != {ex}"
        dense += int((H[i].sum(axis=1)[4:] > 32).sum())
    print(f"{obs_kind}: records {tracks.min()}..{tracks.max()}, dense rows past row 3: {dense}", flush=True)
    assert dense > 0
    srv = Server(device=0, arena_bytes=2 << 30, max_inflight=64, device_loop=True)
    sid = srv.add_system(model, obs)
    for i, p in enumerate(props):
        q = _abi.pcs_proposal.from_buffer_copy(p)
        q.first_video = 0
        assert srv.submit(i, sid, q, f[i:i + 1], n[i:i + 1], 47)
    got = {}
    t0 = time.time()
    while len(got) < len(props):
        for c in srv.poll(256, wait_us=2000):
            assert c.status == 0
            got[c.tag] = (int(c.tracks), int(c.exact))
        assert time.time() - t0 < 120.0, f"only {len(got)} of {len(props)} completions"
    srv.close()
    for i in range(len(props)):
        assert got[i] == (int(tracks[i]), int(exact[i])), i
