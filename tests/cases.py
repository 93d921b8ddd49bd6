"""Workloads shared by tests, smoke() and bench.py.

Physics of the reference's headline configuration: box 200x200x30 um with
periodic boundaries, observation volume 160x120x2 um, two species of 175 nm
and 500 nm diameter (Stokes-Einstein at 295 K, 0.95 mPa s), dt = 0.05 s,
kmin = 3 (test_files/input.xml:9), DE bins over [0, 20] um^2/s.
"""
from __future__ import annotations

import math

import numpy as np

from part_conc_spt_abc_amd import _abi
from part_conc_spt_abc_amd.model import Model, proposal_record, stokes_einstein

D175 = stokes_einstein(175.0)
D500 = stokes_einstein(500.0)


def headline_model(n_frames: int = 1000, n_k: int = 1000, n_de: int = 500, **kw) -> Model:
    args = dict(Lx=200.0, Ly=200.0, Lz=30.0, ax=160.0, ay=120.0, deltat=0.05, kmin=3, n_k=n_k, n_de=n_de,
                ub_de=20.0, n_species=2)
    args.update(kw)
    return Model(**args)


def cpu_model() -> Model:
    """CPU-baseline configuration: 500 frames, H = 500 x 200."""
    return headline_model(n_k=500, n_de=200)


def draw_proposals(model: Model, n: int, rng: np.random.Generator, frames_choices=(1000,), videos: int = 1,
                   c_range=(1e8, 1e10), az: float = 2.0, m=None, first_sim_id: int = 0, max_particles=None):
    """n proposals with total concentration log-uniform on c_range, split
    50/50 between the species; per-video particle counts ~ Poisson(lambda)
    (src/gpu/detail/simulation_impl.inl:343-355). Returns (props, frames, particles)."""
    m = list(m) if m is not None else [D175, D500][: model.n_species]
    props, frames, parts = [], [], []
    for i in range(n):
        ct = 10.0 ** rng.uniform(math.log10(c_range[0]), math.log10(c_range[1]))
        c = [ct / model.n_species] * model.n_species
        lam = ct * model.volume_factor
        f = rng.choice(frames_choices, size=videos)
        p = rng.poisson(lam, size=videos)
        if max_particles is not None:
            p = np.minimum(p, max_particles)
        props.append(proposal_record(model, m, c, [az], first_sim_id + i, len(frames), videos))
        frames.extend(int(x) for x in f)
        parts.extend(int(x) for x in p)
    return props, np.array(frames, np.uint32), np.array(parts, np.uint32)


def fixed_proposal(model: Model, particles, frames, m=None, c=None, az=(2.0,), sim_id=0):
    m = list(m) if m is not None else [D175, D500][: model.n_species]
    c = list(c) if c is not None else [5e8] * model.n_species
    p = proposal_record(model, m, c, list(az), sim_id, 0, len(frames))
    return p, np.asarray(frames, np.uint32), np.asarray(particles, np.uint32)


def split_videos(props, frames, particles):
    """Per-proposal (frames, particles) slices."""
    out = []
    for p in props:
        a, b = p.first_video, p.first_video + p.n_videos
        out.append((frames[a:b], particles[a:b]))
    return out


PCS = _abi
