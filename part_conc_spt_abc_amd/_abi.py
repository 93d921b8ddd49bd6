"""ctypes declarations of include/pcs_abi.h and the loader of libpcs.so.

The library is the only compute path: if it is missing or fails to load,
``load()`` raises instead of falling back to anything else.
"""
from __future__ import annotations

import ctypes as C
import os
import pathlib

PCS_MAX_SPECIES = 8
PCS_MAX_NK = 4096
PCS_MAX_NDE = 4096

PCS_OK = 0
PCS_EINVAL = -1
PCS_ENOMEM = -2
PCS_EHIP = -3
PCS_EAGAIN = -4
PCS_EOVERFLOW = -5
PCS_ENODEV = -6
PCS_ESTATE = -7

PCS_Z_PERIODIC = 0
PCS_Z_REFLECT = 1
PCS_DETECT_SLAB = 0
PCS_DETECT_GAUSS_Z = 1
PCS_DIST_L2SQ = 0
PCS_DIST_L1 = 1
PCS_MODE_PLAIN = 0
PCS_MODE_BATCHED = 1
PCS_MODE_BATCHED_DIRECT = 2
PCS_MODE_PLAIN_STREAMS = 3
PCS_PLAIN_STREAMS = 4
PCS_SERVE_DEVICE_LOOP = 1
PCS_SERVE_LOOP_ONE_WAVE_TILES = 2
PCS_STATIC_SCRATCH_BYTES = 128 << 20
PCS_STATIC_MAX_SIMS = 8192
PCS_STATIC_MAX_VIDEOS = 65536

F8 = C.c_float * PCS_MAX_SPECIES


class pcs_system(C.Structure):
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
        ("dist_kind", C.c_uint32),
    ]


class pcs_proposal(C.Structure):
    _fields_ = [
        ("sim_id", C.c_uint32),
        ("first_video", C.c_uint32),
        ("n_videos", C.c_uint32),
        ("reserved", C.c_uint32),
        ("half_az", F8),
        ("cum_prob", F8),
        ("sigma", F8),
    ]


class pcs_tuning(C.Structure):
    _fields_ = [
        ("particles_per_lane", C.c_uint32),
        ("waves_per_group", C.c_uint32),
        ("groups_per_cu", C.c_uint32),
        ("stage_records", C.c_uint32),
    ]


class pcs_prior(C.Structure):
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


class pcs_server_config(C.Structure):
    _fields_ = [
        ("arena_bytes", C.c_uint64),
        ("max_inflight", C.c_uint32),
        ("max_videos", C.c_uint32),
        ("flags", C.c_uint32),
        ("loop_idle_us", C.c_uint32),
    ]


class pcs_request(C.Structure):
    _fields_ = [
        ("tag", C.c_uint64),
        ("seed", C.c_uint64),
        ("sys_id", C.c_uint32),
        ("n_videos", C.c_uint32),
        ("frames", C.POINTER(C.c_uint32)),
        ("particles", C.POINTER(C.c_uint32)),
        ("prop", pcs_proposal),
    ]


class pcs_completion(C.Structure):
    _fields_ = [
        ("tag", C.c_uint64),
        ("dist", C.c_double),
        ("tracks", C.c_uint32),
        ("status", C.c_int32),
        ("submit_us", C.c_double),
        ("done_us", C.c_double),
        ("exact", C.c_uint64),
    ]


P = C.c_void_p
U32P = C.POINTER(C.c_uint32)
U64P = C.POINTER(C.c_uint64)
F64P = C.POINTER(C.c_double)

# name -> (restype, argtypes); every entry point declared in include/pcs_abi.h.
PROTOTYPES = {
    "pcs_abi_version": (C.c_int, []),
    "pcs_strerror": (C.c_char_p, [C.c_int]),
    "pcs_last_error": (C.c_char_p, []),
    "pcs_device_count": (C.c_int, []),
    "pcs_reference_hist": (C.c_int, [P, P, C.c_uint64, C.c_double, C.c_uint32, C.c_uint32, P, P]),
    "pcs_open": (C.c_int, [C.POINTER(pcs_system), P, C.c_int, C.POINTER(P)]),
    "pcs_close": (C.c_int, [P]),
    "pcs_simulate": (C.c_int, [P, P, C.c_uint32, P, P, C.c_uint32, C.c_uint64, C.c_int, P, P, P]),
    "pcs_simulate_static": (C.c_int, [P, P, C.c_uint32, P, P, C.c_uint32, C.c_uint64, P, P, P]),
    "pcs_batch_create": (C.c_int, [P, C.POINTER(P)]),
    "pcs_batch_load": (C.c_int, [P, P, C.c_uint32, P, P, C.c_uint32]),
    "pcs_batch_run": (C.c_int, [P, C.c_uint64, C.c_int]),
    "pcs_batch_sync": (C.c_int, [P]),
    "pcs_batch_read": (C.c_int, [P, P, P, P]),
    "pcs_batch_read_exact": (C.c_int, [P, P]),
    "pcs_batch_timing": (C.c_int, [P, C.POINTER(C.c_float)]),
    "pcs_batch_overflow_reruns": (C.c_int, [P, U32P]),
    "pcs_batch_arena": (C.c_int, [P, U64P, U64P]),
    "pcs_batch_destroy": (C.c_int, [P]),
    "pcs_batch_propose": (C.c_int, [P, C.POINTER(pcs_prior), P, P, P, C.c_uint32, C.c_uint32, P, C.c_uint32,
                                    C.c_uint32, C.c_uint64, P]),
    "pcs_batch_read_proposals": (C.c_int, [P, P, P]),
    "pcs_batch_set_population": (C.c_int, [P, P, P, C.c_uint32, C.c_uint32]),
    "pcs_batch_pmc_weights": (C.c_int, [P, P, P]),
    "pcs_batch_pmc_weights_async": (C.c_int, [P, P]),
    "pcs_batch_pmc_weights_wait": (C.c_int, [P, P]),
    "pcs_set_tuning": (C.c_int, [P, C.POINTER(pcs_tuning)]),
    "pcs_get_tuning": (C.c_int, [P, C.POINTER(pcs_tuning)]),
    "pcs_set_record_capacity_factor": (C.c_int, [P, C.c_double]),
    "pcs_autotune": (C.c_int, [P, P, C.c_uint32, P, P, C.c_uint32, C.POINTER(pcs_tuning), C.POINTER(C.c_float)]),
    "pcs_selftest_math": (C.c_int, [C.c_int, P, P, P, P, C.c_uint64]),
    "pcs_selftest_copy": (C.c_int, [C.c_int, C.c_uint64, C.POINTER(C.c_float)]),
    "pcs_selftest_sqrt": (C.c_int, [C.c_int, C.c_uint32, C.c_uint32, C.POINTER(C.c_uint64)]),
    "pcs_generate_tracks": (C.c_int, [P, C.POINTER(pcs_proposal), P, P, C.c_uint32, C.c_uint64, P, P, C.c_uint64,
                                       C.POINTER(C.c_uint64)]),
    "pcs_pmc_weights": (C.c_int, [C.c_int, P, C.c_uint32, P, P, C.c_uint32, P, C.c_uint32, P]),
    "pcs_server_create": (C.c_int, [C.c_int, C.POINTER(pcs_server_config), C.POINTER(P)]),
    "pcs_server_add_system": (C.c_int, [P, C.POINTER(pcs_system), P, U32P]),
    "pcs_server_submit": (C.c_int, [P, C.POINTER(pcs_request)]),
    "pcs_server_poll": (C.c_int, [P, C.POINTER(pcs_completion), C.c_uint32, U32P, C.c_uint32]),
    "pcs_server_pending": (C.c_int, [P, U32P]),
    "pcs_server_now_us": (C.c_int, [P, F64P]),
    "pcs_server_set_record_capacity_factor": (C.c_int, [P, C.c_double]),
    "pcs_server_overflow_reruns": (C.c_int, [P, U64P]),
    "pcs_server_group_walked": (C.c_int, [P, U64P]),
    "pcs_server_destroy": (C.c_int, [P]),
}

_LIB = None


def library_path() -> pathlib.Path:
    env = os.environ.get("PCS_LIBRARY")
    if env:
        return pathlib.Path(env)
    return pathlib.Path(__file__).resolve().parent / "lib" / "libpcs.so"


class PcsError(RuntimeError):
    def __init__(self, code: int, where: str, message: str):
        super().__init__(f"{where}: {message} (code {code})")
        self.code = code


def load() -> C.CDLL:
    """Load libpcs.so (built in-tree by __graft_entry__.build()); raise if absent."""
    global _LIB
    if _LIB is not None:
        return _LIB
    path = library_path()
    if not path.exists():
        raise FileNotFoundError(
            f"{path} not found: build the HIP library first (python -c 'import __graft_entry__ as g; g.build()')"
        )
    lib = C.CDLL(str(path))
    for name, (res, args) in PROTOTYPES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = lib
    return lib


def check(code: int, where: str) -> None:
    if code != PCS_OK:
        lib = load()
        msg = lib.pcs_last_error().decode(errors="replace") or lib.pcs_strerror(code).decode()
        raise PcsError(code, where, msg)
