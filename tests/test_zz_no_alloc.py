"""pcs_simulate_static, the entry point that allocates no device memory (its
buffers live in a fixed device array per GPU), against pcs_simulate on the
same proposals: log10 distances bit-equal, track counts and exact sums
equal. Two calls in a row, the second smaller than the first (the array
keeps the first call's contents), then one with record slices forced to
overflow (reruns inside the array); size limits are refused on the host."""
import ctypes as C

import numpy as np
import pytest

import cases
from part_conc_spt_abc_amd import _abi
from part_conc_spt_abc_amd.gpu import Context
from part_conc_spt_abc_amd.model import proposals_array

pytestmark = pytest.mark.gpu


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def simulate_static(ctx, props, frames, parts, seed):
    lib = _abi.load()
    arr = proposals_array(props)
    f = np.ascontiguousarray(frames, np.uint32)
    n = np.ascontiguousarray(parts, np.uint32)
    dist = np.empty(len(props), np.float64)
    tracks = np.empty(len(props), np.uint32)
    exact = np.empty(len(props), np.uint64)
    _abi.check(lib.pcs_simulate_static(ctx.h, C.cast(arr, C.c_void_p), len(props), _p(f), _p(n), len(f), seed,
                                       _p(dist), _p(tracks), _p(exact)), "pcs_simulate_static")
    return dist, tracks, exact


def simulate_alloc(ctx, props, frames, parts, seed):
    """pcs_simulate (batched form) plus the exact sums of the same run."""
    lib = _abi.load()
    arr = proposals_array(props)
    f = np.ascontiguousarray(frames, np.uint32)
    n = np.ascontiguousarray(parts, np.uint32)
    dist = np.empty(len(props), np.float64)
    tracks = np.empty(len(props), np.uint32)
    _abi.check(lib.pcs_simulate(ctx.h, C.cast(arr, C.c_void_p), len(props), _p(f), _p(n), len(f), seed,
                                _abi.PCS_MODE_BATCHED, _p(dist), None, _p(tracks)), "pcs_simulate")
    b = ctx.batch()
    b.load(props, f, n)
    b.run(seed, _abi.PCS_MODE_BATCHED)
    d2, t2, _ = b.read()
    exact = b.read_exact()
    b.close()
    assert np.array_equal(d2.view(np.uint64), dist.view(np.uint64)) and np.array_equal(t2, tracks)
    return dist, tracks, exact


def same(a, b):
    da, ta, ea = a
    db, tb, eb = b
    assert np.array_equal(da.view(np.uint64), db.view(np.uint64))
    assert np.array_equal(ta, tb)
    assert np.array_equal(ea, eb)


def test_static_entry_point_equals_pcs_simulate(gpu_lib):
    model = cases.headline_model()  # H 1000 x 500
    rng = np.random.default_rng(41)
    obs = np.zeros((model.n_k, model.n_de), np.uint32)
    obs[2:80, :150] = rng.integers(0, 5, (78, 150))
    big = cases.draw_proposals(model, 300, rng, frames_choices=(500, 1000, 2000), videos=2, c_range=(1e8, 1e10))
    small = cases.draw_proposals(model, 37, rng, frames_choices=(300, 700), c_range=(1e8, 3e9), first_sim_id=1000)
    with Context(model, obs) as ctx:
        for name, (props, f, n), seed in (("first call, 300 simulations", big, 5), ("second call, 37", small, 6)):
            print(name, flush=True)
            got = simulate_static(ctx, props, f, n, seed)
            assert int(got[1].sum()) > 0
            same(got, simulate_alloc(ctx, props, f, n, seed))
        # record slices far too small: every simulation overflows and is
        # rerun inside the array with an exact slice
        print("forced overflow", flush=True)
        ctx.set_record_capacity_factor(1e-4)
        props, f, n = small
        same(simulate_static(ctx, props, f, n, 8), simulate_alloc(ctx, props, f, n, 8))


def test_static_entry_point_refuses_sizes_beyond_its_limits(gpu_lib):
    model = cases.cpu_model()
    rng = np.random.default_rng(2)
    lib = _abi.load()
    with Context(model) as ctx:
        props, f, n = cases.draw_proposals(model, 3, rng, c_range=(1e8, 1e9))
        arr = proposals_array(props)
        dist = np.empty(_abi.PCS_STATIC_MAX_SIMS + 1, np.float64)
        r = lib.pcs_simulate_static(ctx.h, C.cast(arr, C.c_void_p), _abi.PCS_STATIC_MAX_SIMS + 1, _p(f), _p(n),
                                    len(f), 1, _p(dist), None, None)
        assert r == _abi.PCS_EINVAL and b"at most" in lib.pcs_last_error()
        # videos that do not follow the previous proposal's
        props[1].first_video = 2
        arr = proposals_array(props)
        r = lib.pcs_simulate_static(ctx.h, C.cast(arr, C.c_void_p), 3, _p(f), _p(n), len(f), 1, _p(dist), None, None)
        assert r == _abi.PCS_EINVAL and b"first_video" in lib.pcs_last_error()
