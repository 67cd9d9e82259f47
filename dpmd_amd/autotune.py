"""Autotuner front end (dpmd_autotune and friends in include/dpmd/dpmd.h).

The tuner picks, for a loaded disk batch, the Verlet skin (neighbor cutoff = skin x the
system's smallest diameter, rebuild threshold (cutoff / 2)^2), the cell size (a multiple
``cell_factor`` >= 1 of 2 r_max + cutoff), the workgroup size of the step kernels, the fused
velocity-Verlet switch and the largest hipGraph batch, by timing short device-driven runs
from the current state.  Every candidate grid is checked against the periodic 3x3 stencil's
preconditions (at least 3 x 3 cells, cells no narrower than 2 r_max + cutoff) before it is
timed.  Results persist in a JSON cache keyed by a workload signature.

The reference has no tuner: its cell size, skin and launch shapes are fixed by the caller
(base_particle.hpp:552-592 derives the cell size from the largest diameter and cutoff).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

from . import _lib as L
from ._lib import TuneConfig, TuneTrial, check, lib
from .system import _Particles, _per_system

FIRE, NVE = L.TUNE_FIRE, L.TUNE_NVE


@dataclass
class TuneResult:
    config: dict          # skin, cell_factor, block, max_batch, fused_vv
    rate: float           # particle-steps per second of the chosen configuration
    trials: list          # [(config dict, rate, seconds)] of the trials run (empty on a cache hit)
    signature: str


def cell_grid(Lx: float, Ly: float, r_max: float, cutoff: float, cell_factor: float = 1.0):
    """(nx, ny) of the cell grid for one box; raises DpmdError if the 3x3 stencil's
    preconditions cannot hold (fewer than 3 cells per axis)."""
    nx, ny = C.c_int(), C.c_int()
    check(lib().dpmd_cell_grid(Lx, Ly, r_max, cutoff, cell_factor, C.byref(nx), C.byref(ny)))
    return nx.value, ny.value


def current(P: _Particles) -> dict:
    cfg = TuneConfig()
    P._c(lib().dpmd_tune_get(P.ctx, C.byref(cfg)))
    return cfg.as_dict()


def apply(P: _Particles, config: dict) -> None:
    """Apply a configuration (missing keys keep the context's current values) and rebuild."""
    cur = current(P)
    cur.update(config)
    cfg = TuneConfig(**cur)
    P._c(lib().dpmd_tune_apply(P.ctx, C.byref(cfg)))


def signature(P: _Particles, integrator: int = FIRE, mode: int = L.FORCE_PER_SYSTEM) -> str:
    buf = C.create_string_buffer(1024)
    P._c(lib().dpmd_tune_signature(P.ctx, integrator, mode, buf, len(buf)))
    return buf.value.decode()


def autotune(P: _Particles, integrator: int, dt, mode: int = L.FORCE_PER_SYSTEM, steps: int = 256,
             cache: str | None = None, max_trials: int = 64) -> TuneResult:
    """Tune P for ``integrator`` (FIRE or NVE) at per-system time steps ``dt``; apply and
    return the best configuration.  Positions and velocities are left as they were; FIRE or
    velocity-Verlet state must be initialised again afterwards."""
    S = P.n_systems()
    d = _per_system(dt, S)
    best = TuneConfig()
    rate = C.c_double()
    trials = (TuneTrial * max_trials)()
    n = C.c_int()
    P._c(lib().dpmd_autotune(P.ctx, integrator, d.ptr, mode, steps, cache.encode() if cache else None,
                             C.byref(best), C.byref(rate), trials, max_trials, C.byref(n)))
    rec = [(trials[k].config.as_dict(), trials[k].rate, trials[k].seconds) for k in range(min(n.value, max_trials))]
    return TuneResult(best.as_dict(), rate.value, rec, signature(P, integrator, mode))


def cache_load(path: str, sig: str):
    """(config dict, rate) stored for ``sig`` in the cache at ``path``, or None."""
    cfg, rate, found = TuneConfig(), C.c_double(), C.c_int()
    check(lib().dpmd_tune_cache_load(path.encode(), sig.encode(), C.byref(cfg), C.byref(rate), C.byref(found)))
    return (cfg.as_dict(), rate.value) if found.value else None


def cache_store(path: str, sig: str, config: dict, rate: float) -> None:
    cfg = TuneConfig(**config)
    check(lib().dpmd_tune_cache_store(path.encode(), sig.encode(), C.byref(cfg), rate))
