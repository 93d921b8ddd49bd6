"""Pins the oracle's Philox4x32-10 and FP32 math.

- Philox: the Random123 known-answer vectors (Salmon et al., SC'11,
  kat_vectors for philox4x32_10) and 64 vectors from rocRAND's independent
  implementation (tests/golden/philox_rocrand.json).
- log / sqrt / sincos: every one of the 2^24 inputs the walk can feed them,
  against libm FP64 through numpy, with ulp bounds; plus identities a wrong
  quadrant, sign or swapped pair would break.
- the Gaussian detection draw: its 2^f polynomial against numpy's exp2, and
  its decision against the exact rule u <= exp(-z^2 / (2 s^2)) in FP64.
"""
import json
import math
import pathlib

import numpy as np
import pytest

import oracle_lib as O

GOLDEN = pathlib.Path(__file__).parent / "golden"

# Random123 kat_vectors, philox4x32_10 (counter, key, expected output).
R123_KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


@pytest.mark.parametrize("ctr,key,out", R123_KAT)
def test_philox_random123_kat(ctr, key, out):
    assert tuple(int(x) for x in O.philox(ctr, key)) == out


def test_philox_matches_rocrand():
    vecs = json.loads((GOLDEN / "philox_rocrand.json").read_text())
    assert len(vecs) == 64
    for v in vecs:
        assert [int(x) for x in O.philox(v["ctr"], v["key"])] == v["out"], v


@pytest.fixture(scope="module")
def all_words():
    return (np.arange(1 << 24, dtype=np.uint64) << 8).astype(np.uint32)


def ulps(got, want):
    want32 = want.astype(np.float32)
    sp = np.spacing(np.abs(want32)).astype(np.float64)
    return np.abs(got.astype(np.float64) - want) / sp


def test_radius_all_inputs(all_words):
    r = O.bm_radius(all_words)
    # u = (float)(w | 255) 2^-32 (numpy's uint32 -> float32 cast rounds to
    # nearest even, as the walk's conversion does)
    x = (all_words | np.uint32(0xFF)).astype(np.float32).astype(np.float64)
    assert x.min() == 255.0 and x.max() == 2.0 ** 32
    J = np.arange(1 << 24, dtype=np.float64)
    assert np.array_equal(x, np.where(J < 2 ** 16, 256.0 * J + 255.0, 256.0 * (J + 1.0)))
    want = np.sqrt(-2.0 * np.log(x * 2.0 ** -32))
    # u = 1: -2 ln u is 2^-100 Q(0) with Q(0) = -P0/2 (the tiny term that keeps
    # the square root's input positive), P0 = -0x1.000060p-1; its root, 4.4e-16
    y1 = np.float64(2.0 ** -100) * (-float.fromhex("-0x1.000060p-1") / 2.0)
    assert r[-1] == np.float32(np.sqrt(y1)) and 0.0 < r[-1] < 1e-15
    e = ulps(r[:-1], want[:-1])
    assert e.max() <= 2.5, e.max()  # a degree-6 P: 2.3 ulp at most, 0.45 on average
    assert e.mean() <= 0.5, e.mean()
    # radius is non-increasing in u
    assert np.all(np.diff(r.astype(np.float64)) <= 0)


def test_sincos_all_codes():
    """The Box-Muller angle has 2^16 codes a, theta = 2 pi (a + 1/2) 2^-16:
    table (sin, cos) at 2 pi (i + 1/2) / 256 rotated by the remaining angle."""
    codes = np.arange(1 << 16, dtype=np.uint32)
    s, c = O.sincos_angle(codes << 16)
    th = 2.0 * np.pi * (codes.astype(np.float64) + 0.5) * 2.0 ** -16
    ws, wc = np.sin(th), np.cos(th)
    assert np.abs(s - ws).max() < 1e-7
    assert np.abs(c - wc).max() < 1e-7
    big = np.abs(ws) > 1e-2
    assert ulps(s[big], ws[big]).max() <= 2.0
    big = np.abs(wc) > 1e-2
    assert ulps(c[big], wc[big]).max() <= 2.0
    # quarter- and half-turn symmetries hold exactly (the table is correctly
    # rounded, so its entries are exactly rotated copies of each other)
    q = 1 << 14
    k = np.arange(q)
    assert np.array_equal(s[k + q], c[k])       # sin(t + 1/4) = cos t
    assert np.array_equal(c[k + q], -s[k])      # cos(t + 1/4) = -sin t
    assert np.array_equal(s[k + 2 * q], -s[k])  # half-turn symmetry
    assert np.array_equal(c[k + 2 * q], -c[k])
    norm = s.astype(np.float64) ** 2 + c.astype(np.float64) ** 2
    assert np.abs(norm - 1.0).max() < 4e-7


def test_box_muller_normals_are_standard():
    """Pairs drawn from Philox words through radius/sincos are N(0,1) and
    uncorrelated (moments over 2^20 pairs)."""
    rng = np.random.default_rng(0)
    w = rng.integers(0, 2 ** 32, size=(1 << 20, 2), dtype=np.uint64).astype(np.uint32)
    r = O.bm_radius(w[:, 0]).astype(np.float64)
    s, c = O.sincos_angle(w[:, 1])
    n1, n2 = r * s, r * c
    for n in (n1, n2):
        assert abs(n.mean()) < 5e-3
        assert abs(n.var() - 1.0) < 5e-3
        assert abs((n ** 4).mean() - 3.0) < 3e-2
    assert abs(np.mean(n1 * n2)) < 5e-3


def test_angle_layout():
    """Pair p's 16-bit angle code is the low byte of its radius word above
    byte p % 4 of the chunk's angle word."""
    L = O.lib()
    wr, wa = 0x123456AB, 0xDDCCBBAA
    assert [L.orc_bm_angle16(wr, wa, p) for p in range(4)] == [0xABAA, 0xABBB, 0xABCC, 0xABDD]
    assert L.orc_bm_angle16(0xFFFFFF00, 0xFFFFFFFF, 0) == 0x00FF  # the radius bits never leak in
    assert L.orc_bm_angle16(0x000000FF, 0, 1) == 0xFF00


def test_exp2_polynomial():
    """2^f on [0, 1) within 1.1 ulp of numpy's exp2 at 4096 points spread over
    the interval, and 2^0 = 1 exactly (detection is certain at z = 0)."""
    L = O.lib()
    f = np.linspace(0.0, 1.0, 1 << 22, endpoint=False).astype(np.float32)
    got = np.array([L.orc_exp2_01(float(x)) for x in f[:: 1 << 10]], np.float64)
    want = np.exp2(f[:: 1 << 10].astype(np.float64))
    assert (np.abs(got - want) / np.spacing(want.astype(np.float32))).max() <= 1.1
    assert L.orc_exp2_01(0.0) == 1.0


def test_gauss_detect_decision():
    """The detection draw equals the exact decision u <= exp(-z^2 / (2 s^2)),
    u = (float)(w | 255) 2^-32, computed in FP64, for every (z, w) whose u is
    not within 1e-6 relative of the bound; at z = 0 every w is detected, far
    out (c z^2 beyond the int range of the exponent) none is, and over
    uniform w the detected fraction is exp(-z^2 / (2 s^2))."""
    L = O.lib()
    rng = np.random.default_rng(3)
    for sig in (0.5, 1.0, 2.5):
        c = L.orc_gauss_c(np.float32(sig) * np.float32(sig))
        z = rng.uniform(-4 * sig, 4 * sig, 4000).astype(np.float32)
        w = rng.integers(0, 2 ** 32, 4000, dtype=np.uint64).astype(np.uint32)
        got = np.array([L.orc_gauss_detect(float(a), c, int(b)) for a, b in zip(z, w)], bool)
        u = (w | np.uint32(0xFF)).astype(np.float32).astype(np.float64) * 2.0 ** -32
        bound = np.exp(-z.astype(np.float64) ** 2 / (2.0 * np.float32(sig) ** 2))
        clear = np.abs(u - bound) > 1e-6 * bound
        assert np.array_equal(got[clear], (u <= bound)[clear])
        assert clear.mean() > 0.99
        assert all(L.orc_gauss_detect(0.0, c, int(x)) for x in (0, 0xFFFFFFFF, 0x12345678))
    for tiny in (1e-3, 1e-6, 1e-18):  # c z^2 up to 1e38: the clamped exponent, never detected
        c = L.orc_gauss_c(np.float32(tiny) * np.float32(tiny))
        assert not any(L.orc_gauss_detect(30.0, c, int(x)) for x in (0, 0xFF, 0x12345678))
    c = L.orc_gauss_c(1.0)
    w = rng.integers(0, 2 ** 32, 20000, dtype=np.uint64).astype(np.uint32)
    for zz in (0.5, 1.0, 2.0):
        frac = np.mean([L.orc_gauss_detect(zz, c, int(b)) for b in w])
        p = math.exp(-zz * zz / 2.0)
        assert abs(frac - p) < 5 * math.sqrt(p * (1 - p) / len(w))


@pytest.mark.parametrize("gauss", [False, True])
def test_walk_stream_normals(gauss):
    """The twelve displacements of a chunk, drawn through the walk's stream
    layout (six radius words, angle bytes shared between pairs), are N(0,1)
    and mutually uncorrelated, also between pairs that take their angle
    bytes from the same word."""
    from scipy import stats

    d = O.chunk_normals(seed=0x5eed1234, sim_id=7, video=3, n_particles=8192, n_chunks=8, gauss=gauss)
    x = d.reshape(-1, 12).astype(np.float64)
    n = x.shape[0]
    tol = 5.0 / np.sqrt(n)
    assert np.all(np.abs(x.mean(axis=0)) < tol)
    assert np.all(np.abs(x.var(axis=0) - 1.0) < 5.0 * np.sqrt(2.0 / n))
    assert np.all(np.abs((x ** 4).mean(axis=0) - 3.0) < 5.0 * np.sqrt(96.0 / n))
    corr = np.corrcoef(x, rowvar=False)
    off = corr[~np.eye(12, dtype=bool)]
    assert np.abs(off).max() < tol
    # squares are uncorrelated too (sin/cos of one pair share the radius:
    # their squares correlate through r^2 only as N(0,1) pairs do, zero)
    csq = np.corrcoef(x ** 2, rowvar=False)[~np.eye(12, dtype=bool)]
    assert np.abs(csq).max() < 6.0 / np.sqrt(n)
    assert stats.kstest(x.ravel(), "norm").pvalue > 1e-3
