"""ctypes binding of the CPU oracle (oracle/liboracle.so) for tests.

The oracle is the checker: tests call it to produce expected values, never
as the thing being tested on the GPU path.
"""
from __future__ import annotations

import ctypes as C
import pathlib
import subprocess

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parents[1]
ORACLE_DIR = ROOT / "oracle"
MAX_SPECIES = 8
F8 = C.c_float * MAX_SPECIES
D8 = C.c_double * MAX_SPECIES


class orc_system(C.Structure):
    _fields_ = [
        ("n_species", C.c_uint32),
        ("kmin", C.c_uint32),
        ("n_k", C.c_uint32),
        ("n_de", C.c_uint32),
        ("de_bin_width", C.c_float),
        ("four_dt", C.c_float),
        ("half_l", C.c_float * 3),
        ("half_a", C.c_float * 2),
        ("z_boundary", C.c_uint32),
        ("detect_model", C.c_uint32),
        ("detect_sigma", C.c_float),
    ]


class orc_proposal(C.Structure):
    _fields_ = [
        ("sim_id", C.c_uint32),
        ("n_videos", C.c_uint32),
        ("frames", C.POINTER(C.c_uint32)),
        ("particles", C.POINTER(C.c_uint32)),
        ("half_az", F8),
        ("cum_prob", F8),
        ("sigma", F8),
    ]


class orc_prior(C.Structure):
    _fields_ = [
        ("n_components", C.c_uint32),
        ("n_depths", C.c_uint32),
        ("m_lo", C.c_double),
        ("m_hi", C.c_double),
        ("log10c_lo", C.c_double),
        ("log10c_hi", C.c_double),
        ("az_lo", C.c_double),
        ("az_hi", C.c_double),
        ("deltat", C.c_double),
        ("volume", C.c_double),
    ]


class orc_proposal_f64(C.Structure):
    _fields_ = [("half_az", D8), ("cum_prob", D8), ("sigma", D8)]


_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        so = ORACLE_DIR / "liboracle.so"
        if not so.exists():
            subprocess.run(["make", "-C", str(ORACLE_DIR)], check=True, capture_output=True)
        L = C.CDLL(str(so))
        vp = C.c_void_p
        L.orc_philox4x32_10.argtypes = [vp, vp, vp]
        L.orc_neg2log_u24.argtypes = [C.c_uint32]
        L.orc_neg2log_u24.restype = C.c_float
        L.orc_exp2_01.argtypes = [C.c_float]
        L.orc_exp2_01.restype = C.c_float
        L.orc_gauss_c.argtypes = [C.c_float]
        L.orc_gauss_c.restype = C.c_float
        L.orc_gauss_detect.argtypes = [C.c_float, C.c_float, C.c_uint32]
        L.orc_gauss_detect.restype = C.c_int
        L.orc_bm_radius.argtypes = [C.c_uint32]
        L.orc_bm_radius.restype = C.c_float
        L.orc_bm_radius_n.argtypes = [vp, vp, C.c_uint64]
        L.orc_sincos_angle_n.argtypes = [vp, vp, vp, C.c_uint64]
        L.orc_simulate_f32.argtypes = [C.POINTER(orc_system), C.POINTER(orc_proposal), C.c_uint64, vp, vp]
        L.orc_simulate_f32.restype = C.c_uint64
        L.orc_simulate_f64.argtypes = [C.POINTER(orc_system), C.POINTER(orc_proposal), C.POINTER(orc_proposal_f64),
                                       C.c_double, C.c_double, C.c_uint64, vp, vp]
        L.orc_simulate_f64.restype = C.c_uint64
        L.orc_tracks_f32.argtypes = [C.POINTER(orc_system), C.POINTER(orc_proposal), C.c_uint64, vp, vp,
                                     C.c_uint64]
        L.orc_tracks_f32.restype = C.c_uint64
        L.orc_trace_particle.argtypes = [C.POINTER(orc_system), C.POINTER(orc_proposal),
                                         C.POINTER(orc_proposal_f64), C.c_uint64, C.c_uint32, C.c_uint32, vp, vp]
        L.orc_distance.argtypes = [vp, vp, C.c_uint32, C.c_uint32, C.c_int, C.POINTER(C.c_double)]
        L.orc_distance.restype = C.c_uint64
        L.orc_distance_ref_f32.argtypes = [vp, vp, C.c_uint32, C.c_uint32]
        L.orc_distance_ref_f32.restype = C.c_float
        L.orc_reference_hist.argtypes = [vp, vp, C.c_uint64, C.c_double, C.c_uint32, C.c_uint32, vp, vp]
        L.orc_reference_hist.restype = C.c_uint64
        L.orc_sat_u32.argtypes = [vp, C.c_uint32, C.c_uint32, vp]
        L.orc_bm_angle16.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32]
        L.orc_bm_angle16.restype = C.c_uint32
        L.orc_chunk_normals_n.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_int, vp]
        L.orc_exp10.argtypes = [C.c_double]
        L.orc_exp10.restype = C.c_double
        L.orc_poisson.argtypes = [C.c_double, C.c_uint64, C.c_uint32, C.c_uint32]
        L.orc_poisson.restype = C.c_uint32
        L.orc_propose.argtypes = [C.POINTER(orc_prior), vp, vp, vp, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint64,
                                  vp, vp, vp, vp, vp, vp]
        _LIB = L
    return _LIB


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def philox(ctr, key):
    c = np.asarray(ctr, dtype=np.uint32)
    k = np.asarray(key, dtype=np.uint32)
    o = np.zeros(4, dtype=np.uint32)
    lib().orc_philox4x32_10(_p(c), _p(k), _p(o))
    return o


def bm_radius(words: np.ndarray) -> np.ndarray:
    w = np.ascontiguousarray(words, dtype=np.uint32)
    out = np.empty(len(w), np.float32)
    lib().orc_bm_radius_n(_p(w), _p(out), len(w))
    return out


def sincos_angle(words: np.ndarray):
    """sin and cos of the Box-Muller angle whose 16-bit code is word >> 16."""
    w = np.ascontiguousarray(words, dtype=np.uint32)
    s = np.empty(len(w), np.float32)
    c = np.empty(len(w), np.float32)
    lib().orc_sincos_angle_n(_p(w), _p(s), _p(c), len(w))
    return s, c


def chunk_normals(seed: int, sim_id: int, video: int, n_particles: int, n_chunks: int, gauss: bool = False):
    """Unit displacements of the walk's random stream: (n_particles, n_chunks, 12)."""
    out = np.empty((n_particles, n_chunks, 12), np.float32)
    lib().orc_chunk_normals_n(seed, sim_id, video, n_particles, n_chunks, int(gauss), _p(out))
    return out


def system_from_abi(sys) -> orc_system:
    """Oracle system with the same FP32 values as a pcs_system."""
    o = orc_system()
    for f in ("n_species", "kmin", "n_k", "n_de", "de_bin_width", "four_dt", "z_boundary", "detect_model",
              "detect_sigma"):
        setattr(o, f, getattr(sys, f))
    for i in range(3):
        o.half_l[i] = sys.half_l[i]
    for i in range(2):
        o.half_a[i] = sys.half_a[i]
    return o


class OracleProposal:
    """Keeps the frame/particle arrays alive next to the struct."""

    def __init__(self, prop, frames, particles):
        self.frames = np.ascontiguousarray(frames, dtype=np.uint32)
        self.particles = np.ascontiguousarray(particles, dtype=np.uint32)
        o = orc_proposal()
        o.sim_id = prop.sim_id
        o.n_videos = len(self.frames)
        o.frames = self.frames.ctypes.data_as(C.POINTER(C.c_uint32))
        o.particles = self.particles.ctypes.data_as(C.POINTER(C.c_uint32))
        for i in range(MAX_SPECIES):
            o.half_az[i] = prop.half_az[i]
            o.cum_prob[i] = prop.cum_prob[i]
            o.sigma[i] = prop.sigma[i]
        self.s = o


def simulate_f32(osys: orc_system, oprop: OracleProposal, seed: int):
    H = np.zeros((osys.n_k, osys.n_de), dtype=np.uint32)
    dropped = np.zeros(1, dtype=np.uint64)
    total = lib().orc_simulate_f32(C.byref(osys), C.byref(oprop.s), seed, _p(H), _p(dropped))
    return H, int(total), int(dropped[0])


def simulate_f64(osys: orc_system, oprop: OracleProposal, pd: orc_proposal_f64, ub_de: float, deltat: float,
                 seed: int):
    H = np.zeros((osys.n_k, osys.n_de), dtype=np.uint32)
    dropped = np.zeros(1, dtype=np.uint64)
    total = lib().orc_simulate_f64(C.byref(osys), C.byref(oprop.s), C.byref(pd), ub_de, deltat, seed, _p(H),
                                   _p(dropped))
    return H, int(total), int(dropped[0])


def tracks_f32(osys: orc_system, oprop: OracleProposal, seed: int, cap: int = 1 << 22):
    K = np.zeros(cap, dtype=np.uint32)
    DE = np.zeros(cap, dtype=np.float32)
    n = lib().orc_tracks_f32(C.byref(osys), C.byref(oprop.s), seed, _p(K), _p(DE), cap)
    n = min(int(n), cap)
    return K[:n].copy(), DE[:n].copy()


def trace_particle(osys, oprop, pd, seed, video, particle):
    F = int(oprop.frames[video])
    x32 = np.zeros((F, 3), np.float32)
    x64 = np.zeros((F, 3), np.float64)
    lib().orc_trace_particle(C.byref(osys), C.byref(oprop.s), C.byref(pd), seed, video, particle, _p(x32), _p(x64))
    return x32, x64


def distance(H: np.ndarray, Hobs: np.ndarray, kind: int = 0):
    H = np.ascontiguousarray(H, dtype=np.uint32)
    Hobs = np.ascontiguousarray(Hobs, dtype=np.uint32)
    lg = C.c_double(0.0)
    v = lib().orc_distance(_p(H), _p(Hobs), H.shape[0], H.shape[1], kind, C.byref(lg))
    return int(v), float(lg.value)


def distance_ref_f32(H, Hobs) -> float:
    H = np.ascontiguousarray(H, dtype=np.uint32)
    Hobs = np.ascontiguousarray(Hobs, dtype=np.uint32)
    return float(lib().orc_distance_ref_f32(_p(H), _p(Hobs), H.shape[0], H.shape[1]))


def reference_hist(K, DE, ub_de, n_k, n_de):
    K = np.ascontiguousarray(K, dtype=np.uint64)
    DE = np.ascontiguousarray(DE, dtype=np.float64)
    H = np.zeros((n_k, n_de), dtype=np.uint32)
    dropped = np.zeros(1, dtype=np.uint64)
    n = lib().orc_reference_hist(_p(K), _p(DE), len(K), ub_de, n_k, n_de, _p(H), _p(dropped))
    return H, int(n), int(dropped[0])


def exp10(x: float) -> float:
    return float(lib().orc_exp10(float(x)))


def poisson(lam: float, seed: int, trial: int, video: int = 0) -> int:
    return int(lib().orc_poisson(float(lam), seed, trial, video))


def propose(prior: orc_prior, theta, cum_w, tau, n_trials: int, J: int, seed: int):
    """Proposal step: returns dict(theta (P, n_trials), half_az / cum_prob /
    sigma (n_trials, 8), particles (n_trials, J), parent (n_trials,))."""
    theta = np.ascontiguousarray(theta, dtype=np.float64)
    cum_w = np.ascontiguousarray(cum_w, dtype=np.float64)
    tau = np.ascontiguousarray(tau, dtype=np.float64)
    P, n_pop = theta.shape
    out = dict(theta=np.zeros((P, n_trials)), half_az=np.zeros((n_trials, MAX_SPECIES), np.float32),
               cum_prob=np.zeros((n_trials, MAX_SPECIES), np.float32),
               sigma=np.zeros((n_trials, MAX_SPECIES), np.float32),
               particles=np.zeros((n_trials, J), np.uint32), parent=np.zeros(n_trials, np.uint32))
    lib().orc_propose(C.byref(prior), _p(theta), _p(cum_w), _p(tau), n_pop, n_trials, J, seed, _p(out["theta"]),
                      _p(out["half_az"]), _p(out["cum_prob"]), _p(out["sigma"]), _p(out["particles"]),
                      _p(out["parent"]))
    return out
