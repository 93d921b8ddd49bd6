"""Python handles over the C ABI (include/pcs_abi.h): Context, Batch, Server.

Every call goes to libpcs.so on the GPU; there is no other compute path.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional, Sequence

import numpy as np

from . import _abi
from .model import Model, proposals_array


def _ptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def device_count() -> int:
    return int(_abi.load().pcs_device_count())


def reference_hist(K: np.ndarray, DE: np.ndarray, ub_de: float, n_k: int, n_de: int):
    """Observed histogram from track lists (host); returns (hist, dropped)."""
    lib = _abi.load()
    K = np.ascontiguousarray(K, dtype=np.uint64)
    DE = np.ascontiguousarray(DE, dtype=np.float64)
    H = np.zeros((n_k, n_de), dtype=np.uint32)
    dropped = C.c_uint64(0)
    _abi.check(lib.pcs_reference_hist(_ptr(K), _ptr(DE), len(K), ub_de, n_k, n_de, _ptr(H), C.byref(dropped)),
               "pcs_reference_hist")
    return H, int(dropped.value)


class Context:
    """One model and its observed histogram on one GPU (pcs_open)."""

    def __init__(self, model: Model, obs_hist: Optional[np.ndarray] = None, device: int = 0):
        self.lib = _abi.load()
        self.model = model
        self.sys = model.to_abi()
        if obs_hist is None:
            obs_hist = np.zeros((model.n_k, model.n_de), dtype=np.uint32)
        obs = np.ascontiguousarray(obs_hist, dtype=np.uint32)
        if obs.shape != (model.n_k, model.n_de):
            raise ValueError(f"observed histogram shape {obs.shape} != {(model.n_k, model.n_de)}")
        h = C.c_void_p()
        _abi.check(self.lib.pcs_open(C.byref(self.sys), _ptr(obs), device, C.byref(h)), "pcs_open")
        self.h = h

    def close(self):
        if self.h:
            self.lib.pcs_close(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_tuning(self, particles_per_lane=0, waves_per_group=0, groups_per_cu=0, stage_records=0):
        t = _abi.pcs_tuning(particles_per_lane, waves_per_group, groups_per_cu, stage_records)
        _abi.check(self.lib.pcs_set_tuning(self.h, C.byref(t)), "pcs_set_tuning")

    def get_tuning(self) -> dict:
        t = _abi.pcs_tuning()
        _abi.check(self.lib.pcs_get_tuning(self.h, C.byref(t)), "pcs_get_tuning")
        return {f: int(getattr(t, f)) for f, _ in t._fields_}

    def autotune(self, props, frames, particles):
        """Pick the fastest launch shape on a calibration batch; returns
        (tuning dict, walk ms)."""
        arr = proposals_array(props)
        fr = np.ascontiguousarray(frames, dtype=np.uint32)
        pa = np.ascontiguousarray(particles, dtype=np.uint32)
        t = _abi.pcs_tuning()
        ms = C.c_float(0.0)
        _abi.check(self.lib.pcs_autotune(self.h, C.cast(arr, C.c_void_p), len(props), _ptr(fr), _ptr(pa), len(fr),
                                         C.byref(t), C.byref(ms)), "pcs_autotune")
        return {f: int(getattr(t, f)) for f, _ in t._fields_}, float(ms.value)

    def set_record_capacity_factor(self, f: float):
        _abi.check(self.lib.pcs_set_record_capacity_factor(self.h, f), "pcs_set_record_capacity_factor")

    def batch(self) -> "Batch":
        return Batch(self)


class Batch:
    """Proposals resident in HBM (pcs_batch_*)."""

    def __init__(self, ctx: Context):
        self.ctx = ctx
        self.lib = ctx.lib
        h = C.c_void_p()
        _abi.check(self.lib.pcs_batch_create(ctx.h, C.byref(h)), "pcs_batch_create")
        self.h = h
        self.n = 0

    def load(self, props: Sequence[_abi.pcs_proposal], frames: np.ndarray, particles: np.ndarray):
        arr = proposals_array(props)
        self._frames = np.ascontiguousarray(frames, dtype=np.uint32)
        self._particles = np.ascontiguousarray(particles, dtype=np.uint32)
        _abi.check(self.lib.pcs_batch_load(self.h, C.cast(arr, C.c_void_p), len(props), _ptr(self._frames),
                                           _ptr(self._particles), len(self._frames)), "pcs_batch_load")
        self.n = len(props)

    def set_population(self, theta: np.ndarray, w: np.ndarray):
        """Keep a population (P x n_pop parameters, weights) resident in the
        batch (pcs_batch_set_population)."""
        th = np.ascontiguousarray(theta, dtype=np.float64)
        ww = np.ascontiguousarray(w, dtype=np.float64)
        _abi.check(self.lib.pcs_batch_set_population(self.h, _ptr(th), _ptr(ww), th.shape[1], th.shape[0]),
                   "pcs_batch_set_population")

    def propose(self, prior: _abi.pcs_prior, theta, cum_w, tau: np.ndarray, n_trials: int,
                frames: np.ndarray, seed: int, first_sim_id: int = 0, want_theta: bool = True):
        """Proposal step of an ABC-PMC round on the GPU (pcs_batch_propose):
        theta (P x n_pop) the population, cum_w its cumulative weights (both
        None: the resident population), tau the displacement scales; frames
        (n_trials x J) the trials' video lengths. Returns the proposed
        parameters (P x n_trials) or None."""
        th = None if theta is None else np.ascontiguousarray(theta, dtype=np.float64)
        cw = None if cum_w is None else np.ascontiguousarray(cum_w, dtype=np.float64)
        ta = np.ascontiguousarray(tau, dtype=np.float64)
        fr = np.ascontiguousarray(frames, dtype=np.uint32).reshape(n_trials, -1)
        out = np.empty((len(ta), n_trials), np.float64) if want_theta else None
        n_pop = 0 if th is None else th.shape[1]
        _abi.check(self.lib.pcs_batch_propose(self.h, C.byref(prior), _ptr(th), _ptr(cw), _ptr(ta), n_pop,
                                              n_trials, _ptr(fr), fr.shape[1], first_sim_id, C.c_uint64(seed),
                                              _ptr(out)), "pcs_batch_propose")
        self._frames = fr.ravel()
        self.n = n_trials
        return out

    def pmc_weights(self, tau: np.ndarray) -> np.ndarray:
        """pmc-standard weights of the proposed trials against the resident
        population, on the GPU (pcs_batch_pmc_weights)."""
        ta = np.ascontiguousarray(tau, dtype=np.float64)
        out = np.empty(self.n, np.float64)
        _abi.check(self.lib.pcs_batch_pmc_weights(self.h, _ptr(ta), _ptr(out)), "pcs_batch_pmc_weights")
        return out

    def pmc_weights_async(self, tau: np.ndarray) -> None:
        """Enqueue the weights on the batch's weights stream
        (pcs_batch_pmc_weights_async); collect them with pmc_weights_wait."""
        ta = np.ascontiguousarray(tau, dtype=np.float64)
        _abi.check(self.lib.pcs_batch_pmc_weights_async(self.h, _ptr(ta)), "pcs_batch_pmc_weights_async")

    def pmc_weights_wait(self) -> np.ndarray:
        out = np.empty(self.n, np.float64)
        _abi.check(self.lib.pcs_batch_pmc_weights_wait(self.h, _ptr(out)), "pcs_batch_pmc_weights_wait")
        return out

    def read_proposals(self):
        """(pcs_proposal records, video particle counts) of the loaded or
        proposed batch."""
        recs = (_abi.pcs_proposal * self.n)()
        parts = np.empty(len(self._frames), np.uint32)
        _abi.check(self.lib.pcs_batch_read_proposals(self.h, C.cast(recs, C.c_void_p), _ptr(parts)),
                   "pcs_batch_read_proposals")
        return list(recs), parts

    def run(self, seed: int, mode: int = _abi.PCS_MODE_BATCHED):
        _abi.check(self.lib.pcs_batch_run(self.h, C.c_uint64(seed), mode), "pcs_batch_run")

    def sync(self):
        _abi.check(self.lib.pcs_batch_sync(self.h), "pcs_batch_sync")

    def read(self, hist: bool = False):
        dist = np.empty(self.n, dtype=np.float64)
        tracks = np.empty(self.n, dtype=np.uint32)
        H = None
        if hist:
            m = self.ctx.model
            H = np.empty((self.n, m.n_k, m.n_de), dtype=np.uint32)
        _abi.check(self.lib.pcs_batch_read(self.h, _ptr(dist), _ptr(H), _ptr(tracks)), "pcs_batch_read")
        return dist, tracks, H

    def read_exact(self) -> np.ndarray:
        out = np.empty(self.n, dtype=np.uint64)
        _abi.check(self.lib.pcs_batch_read_exact(self.h, _ptr(out)), "pcs_batch_read_exact")
        return out

    def timing(self):
        t = (C.c_float * 3)()
        _abi.check(self.lib.pcs_batch_timing(self.h, t), "pcs_batch_timing")
        return float(t[0]), float(t[1]), float(t[2])

    def arena(self):
        """(slice capacity in records, device bytes of the record arenas)."""
        cap, nb = C.c_uint64(0), C.c_uint64(0)
        _abi.check(self.lib.pcs_batch_arena(self.h, C.byref(cap), C.byref(nb)), "pcs_batch_arena")
        return int(cap.value), int(nb.value)

    def overflow_reruns(self) -> int:
        n = C.c_uint32(0)
        _abi.check(self.lib.pcs_batch_overflow_reruns(self.h, C.byref(n)), "pcs_batch_overflow_reruns")
        return int(n.value)

    def close(self):
        if self.h:
            self.lib.pcs_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def pmc_weights(theta_new: np.ndarray, theta_old: np.ndarray, w_old: np.ndarray, tau: np.ndarray,
                device: int = 0) -> np.ndarray:
    """pmc-standard importance weights on the GPU (pcs_pmc_weights): rows are
    parameters, columns members; returns the normalised new weights."""
    lib = _abi.load()
    xn = np.ascontiguousarray(theta_new, dtype=np.float64)
    xo = np.ascontiguousarray(theta_old, dtype=np.float64)
    wo = np.ascontiguousarray(w_old, dtype=np.float64)
    ta = np.ascontiguousarray(tau, dtype=np.float64)
    out = np.empty(xn.shape[1], np.float64)
    _abi.check(lib.pcs_pmc_weights(device, _ptr(xn), xn.shape[1], _ptr(xo), _ptr(wo), xo.shape[1], _ptr(ta),
                                   len(ta), _ptr(out)), "pcs_pmc_weights")
    return out


def selftest_math(words: np.ndarray, device: int = 0):
    lib = _abi.load()
    w = np.ascontiguousarray(words, dtype=np.uint32)
    r = np.empty(len(w), np.float32)
    s = np.empty(len(w), np.float32)
    c = np.empty(len(w), np.float32)
    _abi.check(lib.pcs_selftest_math(device, _ptr(w), _ptr(r), _ptr(s), _ptr(c), len(w)), "pcs_selftest_math")
    return r, s, c


def selftest_sqrt(lo_bits: int, hi_bits: int, device: int = 0) -> int:
    """Number of floats with bit patterns in [lo_bits, hi_bits) on which the
    walk's square root is not correctly rounded (checked on the GPU)."""
    lib = _abi.load()
    n = C.c_uint64(0)
    _abi.check(lib.pcs_selftest_sqrt(device, lo_bits, hi_bits, C.byref(n)), "pcs_selftest_sqrt")
    return int(n.value)


class Server:
    """Continuous-batching serving loop (pcs_server_*). device_loop=True
    selects the device-resident form (one persistent kernel fed from a ring in
    pinned host memory) instead of host-driven launch groups; with it,
    one_wave_tiles=True walks every tile with one wave (a small request's
    tiles are otherwise walked by a whole group each)."""

    def __init__(self, device: int = 0, arena_bytes: int = 0, max_inflight: int = 0, max_videos: int = 0,
                 device_loop: bool = False, loop_idle_us: int = 0, one_wave_tiles: bool = False):
        self.lib = _abi.load()
        flags = _abi.PCS_SERVE_DEVICE_LOOP if device_loop else 0
        if one_wave_tiles:
            flags |= _abi.PCS_SERVE_LOOP_ONE_WAVE_TILES
        cfg = _abi.pcs_server_config(arena_bytes, max_inflight, max_videos, flags, loop_idle_us)
        h = C.c_void_p()
        _abi.check(self.lib.pcs_server_create(device, C.byref(cfg), C.byref(h)), "pcs_server_create")
        self.h = h

    def add_system(self, model: Model, obs_hist: Optional[np.ndarray] = None) -> int:
        sys = model.to_abi()
        if obs_hist is None:
            obs_hist = np.zeros((model.n_k, model.n_de), dtype=np.uint32)
        obs = np.ascontiguousarray(obs_hist, dtype=np.uint32)
        sid = C.c_uint32(0)
        _abi.check(self.lib.pcs_server_add_system(self.h, C.byref(sys), _ptr(obs), C.byref(sid)),
                   "pcs_server_add_system")
        return int(sid.value)

    def submit(self, tag: int, sys_id: int, prop: _abi.pcs_proposal, frames, particles, seed: int) -> bool:
        """Returns False when the queue is full (PCS_EAGAIN)."""
        f = np.ascontiguousarray(frames, dtype=np.uint32)
        p = np.ascontiguousarray(particles, dtype=np.uint32)
        req = _abi.pcs_request()
        req.tag = tag
        req.seed = seed
        req.sys_id = sys_id
        req.n_videos = len(f)
        req.frames = f.ctypes.data_as(C.POINTER(C.c_uint32))
        req.particles = p.ctypes.data_as(C.POINTER(C.c_uint32))
        req.prop = prop
        r = self.lib.pcs_server_submit(self.h, C.byref(req))
        if r == _abi.PCS_EAGAIN:
            return False
        _abi.check(r, "pcs_server_submit")
        return True

    def poll(self, max_n: int = 1024, wait_us: int = 0):
        buf = getattr(self, "_buf", None)
        if buf is None or len(buf) < max_n:
            buf = self._buf = (_abi.pcs_completion * max_n)()
        n = C.c_uint32(0)
        _abi.check(self.lib.pcs_server_poll(self.h, buf, max_n, C.byref(n), wait_us), "pcs_server_poll")
        return [buf[i] for i in range(n.value)]

    def pending(self) -> int:
        n = C.c_uint32(0)
        _abi.check(self.lib.pcs_server_pending(self.h, C.byref(n)), "pcs_server_pending")
        return int(n.value)

    def set_record_capacity_factor(self, f: float):
        _abi.check(self.lib.pcs_server_set_record_capacity_factor(self.h, f), "pcs_server_set_record_capacity_factor")

    def overflow_reruns(self) -> int:
        n = C.c_uint64(0)
        _abi.check(self.lib.pcs_server_overflow_reruns(self.h, C.byref(n)), "pcs_server_overflow_reruns")
        return int(n.value)

    def group_walked(self) -> int:
        """Completed requests whose tiles were each walked by a whole group."""
        n = C.c_uint64(0)
        _abi.check(self.lib.pcs_server_group_walked(self.h, C.byref(n)), "pcs_server_group_walked")
        return int(n.value)

    def now_us(self) -> float:
        """The server's clock (us since creation), the clock of submit_us/done_us."""
        t = C.c_double(0.0)
        _abi.check(self.lib.pcs_server_now_us(self.h, C.byref(t)), "pcs_server_now_us")
        return float(t.value)

    def close(self):
        if self.h:
            self.lib.pcs_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
