"""Model description and per-proposal precompute, mirroring the reference.

- ``Model`` holds what the reference's SimulateSetup holds
  (src/gpu/cusim/kernel-simulate.cuh:13-38, filled at
  src/gpu/detail/simulation_impl.inl:568-579): box, detection extents, time
  step, histogram shape.
- ``proposal_record`` is the per-trial precompute of
  src/gpu/detail/simulation_impl.inl:306-341: half detection depths,
  cumulative abundances with the last entry forced to 2, and random-walk
  steps sqrt(2 m dt).
"""
from __future__ import annotations

import dataclasses
import math
from typing import Sequence

import numpy as np

from . import _abi

BOLTZMANN = 1.380649e-23  # J/K


def stokes_einstein(diameter_nm: float, temperature_k: float = 295.0, viscosity_pa_s: float = 0.95e-3) -> float:
    """Diffusion coefficient in um^2/s of a sphere of the given diameter."""
    r = 0.5 * diameter_nm * 1e-9
    d = BOLTZMANN * temperature_k / (6.0 * math.pi * viscosity_pa_s * r)  # m^2/s
    return d * 1e12


@dataclasses.dataclass
class Model:
    """Physical model shared by every simulation of an ABC run (lengths in um)."""

    Lx: float
    Ly: float
    Lz: float
    ax: float
    ay: float
    deltat: float
    kmin: int
    n_k: int
    n_de: int
    ub_de: float
    n_species: int
    z_boundary: int = _abi.PCS_Z_PERIODIC
    detect_model: int = _abi.PCS_DETECT_SLAB
    detect_sigma: float = 0.0
    dist_kind: int = _abi.PCS_DIST_L2SQ

    @property
    def volume_factor(self) -> float:
        """Expected particles per video per (particle/mL): Lx Ly Lz / 1e12
        (src/cpu/simulate.jl:23, src/gpu/detail/simulation_impl.inl:515)."""
        return self.Lx * self.Ly * self.Lz / 1e12

    @property
    def de_bin_width(self) -> float:
        return self.ub_de / self.n_de

    def to_abi(self) -> _abi.pcs_system:
        s = _abi.pcs_system()
        s.n_species = self.n_species
        s.kmin = self.kmin
        s.n_k = self.n_k
        s.n_de = self.n_de
        s.de_bin_width = np.float32(self.ub_de / self.n_de)
        s.four_dt = np.float32(4.0) * np.float32(self.deltat)
        s.half_l[0] = np.float32(0.5) * np.float32(self.Lx)
        s.half_l[1] = np.float32(0.5) * np.float32(self.Ly)
        s.half_l[2] = np.float32(0.5) * np.float32(self.Lz)
        s.half_a[0] = np.float32(0.5) * np.float32(self.ax)
        s.half_a[1] = np.float32(0.5) * np.float32(self.ay)
        s.z_boundary = self.z_boundary
        s.detect_model = self.detect_model
        s.detect_sigma = self.detect_sigma
        s.dist_kind = self.dist_kind
        return s


def proposal_record(model: Model, m: Sequence[float], c: Sequence[float], az: Sequence[float], sim_id: int,
                    first_video: int, n_videos: int) -> _abi.pcs_proposal:
    """Per-simulation parameters from diffusion coefficients m (um^2/s),
    concentrations c (particles/mL) and detection depths az (um); az may have
    one entry (fixed depth) or one per species."""
    C = model.n_species
    if len(m) != C or len(c) != C:
        raise ValueError("m and c need one entry per species")
    p = _abi.pcs_proposal()
    p.sim_id = sim_id
    p.first_video = first_video
    p.n_videos = n_videos
    p.reserved = 0
    csum = float(sum(c))
    acc = 0.0
    for i in range(C):
        a = az[i] if len(az) > 1 else az[0]
        p.half_az[i] = np.float32(0.5) * np.float32(a)
        acc += c[i] / csum if csum > 0 else (1.0 if i == C - 1 else 0.0)
        p.cum_prob[i] = np.float32(acc)
        p.sigma[i] = np.float32(math.sqrt(2.0 * m[i] * model.deltat))
    # src/gpu/detail/simulation_impl.inl:331-335: last cumulative weight = 2.
    p.cum_prob[C - 1] = 2.0
    return p


def proposals_array(records: Sequence[_abi.pcs_proposal]):
    arr = (_abi.pcs_proposal * len(records))()
    for i, r in enumerate(records):
        arr[i] = r
    return arr


def prior_record(model: Model, m_bounds, log10c_bounds, az_bounds, fixed_depth: bool = True) -> _abi.pcs_prior:
    """Prior of the device proposal step for a model: parameter rows m_1..m_C,
    log10 c_1..log10 c_C, az (one shared depth when fixed, else one per
    component)."""
    C = model.n_species
    return _abi.pcs_prior(C, 1 if fixed_depth else C, float(m_bounds[0]), float(m_bounds[1]),
                          float(log10c_bounds[0]), float(log10c_bounds[1]), float(az_bounds[0]),
                          float(az_bounds[1]), float(model.deltat), float(model.volume_factor))
