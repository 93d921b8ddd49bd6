"""MI355X-native simulation hot path for ABC-PMC particle concentration
estimation: Brownian walk -> detection -> (K, DE) track histogram ->
cumulative-histogram distance, as HIP kernels for gfx950 behind a C ABI
(include/pcs_abi.h, built into part_conc_spt_abc_amd/lib/libpcs.so).
"""
from ._abi import (  # noqa: F401
    PCS_DETECT_GAUSS_Z,
    PCS_DETECT_SLAB,
    PCS_DIST_L1,
    PCS_DIST_L2SQ,
    PCS_MODE_BATCHED,
    PCS_MODE_BATCHED_DIRECT,
    PCS_MODE_PLAIN,
    PCS_MODE_PLAIN_STREAMS,
    PCS_Z_PERIODIC,
    PCS_Z_REFLECT,
    PcsError,
    load,
)
from .model import Model, proposal_record, stokes_einstein  # noqa: F401

__all__ = [
    "Model",
    "proposal_record",
    "stokes_einstein",
    "load",
    "PcsError",
]
