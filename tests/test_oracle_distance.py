"""Pins the oracle's cumulative-histogram distance (src/cpu/distance.jl:1-8,
src/gpu/cusim/kernel-distance.inl:81-241) and the observed-histogram
builder (src/gpu/detail/simulation_impl.inl:605-616) against textbook numpy
and the reference's own test_files fixture.
"""
import hashlib
import math
import json
import pathlib
import re

import numpy as np
import pytest

import oracle_lib as O

GOLDEN = pathlib.Path(__file__).parent / "golden"


def textbook(H, R, kind):
    S = np.cumsum(np.cumsum(H.astype(np.int64), axis=0), axis=1)
    T = np.cumsum(np.cumsum(R.astype(np.int64), axis=0), axis=1)
    d = S - T
    return int((d * d).sum()) if kind == 0 else int(np.abs(d).sum())


@pytest.mark.parametrize("shape", [(1, 1), (7, 3), (50, 40), (200, 120)])
@pytest.mark.parametrize("kind", [0, 1])
def test_distance_against_numpy_cumsum(shape, kind):
    rng = np.random.default_rng(shape[0] * 31 + shape[1] + kind)
    H = (rng.random(shape) < 0.1) * rng.integers(0, 20, shape)
    R = (rng.random(shape) < 0.1) * rng.integers(0, 20, shape)
    v, lg = O.distance(H.astype(np.uint32), R.astype(np.uint32), kind)
    want = textbook(H, R, kind)
    assert v == want
    if want:
        assert lg == pytest.approx(np.log10(want), rel=1e-15)
    assert O.distance(H.astype(np.uint32), H.astype(np.uint32), kind)[0] == 0


def test_distance_hand_example():
    # one track at (row 0, bin 0) vs one at (row 1, bin 1) in a 2x2 matrix:
    # S = [[1,1],[1,1]], R = [[0,0],[0,1]] -> squares 1+1+1+0 = 3
    H = np.array([[1, 0], [0, 0]], np.uint32)
    R = np.array([[0, 0], [0, 1]], np.uint32)
    assert O.distance(H, R, 0)[0] == 3
    assert O.distance(H, R, 1)[0] == 3


def test_reference_float_form_close_to_exact():
    rng = np.random.default_rng(2)
    H = rng.integers(0, 5, (300, 200)).astype(np.uint32)
    R = rng.integers(0, 5, (300, 200)).astype(np.uint32)
    exact = O.distance(H, R, 0)[0]
    f32 = O.distance_ref_f32(H, R)
    assert abs(f32 - exact) / exact < 1e-4


def parse_data_xml(path):
    txt = pathlib.Path(path).read_text()

    def field(name):
        return re.search(rf"<{name}>(.*?)</{name}>", txt, re.S).group(1).strip()

    K = np.array([int(x) for x in field("K").split(",")], np.uint64)
    DE = np.array([float(x) for x in field("DE").split(",")], np.float64)
    return K, DE


def test_observed_histogram_from_test_files():
    golden = json.loads((GOLDEN / "ref_xml_parse.json").read_text())
    K, DE = parse_data_xml(GOLDEN / "test_files" / "data_20171018_175nm_500nm.xml")
    # the reference loader's parse of the same file (lengths, sums, digests)
    assert len(K) == golden["K"]["n"] and float(K.sum()) == golden["K"]["sum"]
    assert hashlib.sha256(",".join("%.17g" % v for v in K).encode()).hexdigest() == golden["K"]["sha256"]
    assert hashlib.sha256(",".join("%.17g" % v for v in DE).encode()).hexdigest() == golden["DE"]["sha256"]
    ub_de, n_de = golden["ub_de"], golden["number_of_de_bins"]
    H, counted, dropped = O.reference_hist(K, DE, ub_de, 1000, n_de)
    # 38,302 tracks, 1,601 with DE >= ub_de = 5.0 are skipped
    assert counted == 36701 and dropped == 0 and H.sum() == 36701
    rows = np.nonzero(H.sum(axis=1))[0]
    assert rows.max() == 211 - 1 and rows.min() == 3 - 1
    # summed-area corner = number of binned tracks
    S = np.zeros_like(H)
    O.lib().orc_sat_u32(H.ctypes.data, 1000, n_de, S.ctypes.data)
    assert S[-1, -1] == 36701


@pytest.mark.parametrize("kind", [0, 1])
def test_distance_sum_past_two_to_the_64(kind):
    """A sum of squares past 2^64 (observed counts near 2^32 in every
    cumulative cell): the exact value saturates at 2^64 - 1 and the log10
    distance is that of the exact (Python integer) sum."""
    rng = np.random.default_rng(5)
    shape = (30, 20)
    R = np.zeros(shape, np.uint32)
    R[0, 0] = 4_000_000_000
    R[3, 5] = 200_000_000
    H = ((rng.random(shape) < 0.2) * rng.integers(0, 50, shape)).astype(np.uint32)
    S = np.cumsum(np.cumsum(H.astype(object), axis=0), axis=1)
    T = np.cumsum(np.cumsum(R.astype(object), axis=0), axis=1)
    d = S - T
    want = int(sum(int(x) * int(x) for x in d.ravel())) if kind == 0 else int(sum(abs(int(x)) for x in d.ravel()))
    v, lg = O.distance(H, R, kind)
    if kind == 0:
        assert want > 2 ** 64 and v == 2 ** 64 - 1
    else:
        assert v == want
    assert lg == pytest.approx(math.log10(want), rel=1e-15)
