"""Many independent systems advanced by one persistent launch per step
(libnbody_amd nb_batch_*).

    b = nbody_amd.Batch([params_0, params_1, ...], pos, vel)   # concatenated arrays
    b.autotune(3)          # project-written autotuner: times every launch shape
    b.step(200)            # persistent kernel, replayed from a hipGraph
"""
from __future__ import annotations

import ctypes
from typing import Sequence

import numpy as np

from ._lib import NB_ERR_INVALID_ARG, NbParams, check, load


def _p(a: np.ndarray) -> ctypes.c_void_p:
    if a.dtype != np.float32 or not a.flags.c_contiguous:
        raise ValueError("expected a contiguous float32 array")
    return ctypes.c_void_p(a.ctypes.data)


def batch_config_count() -> int:
    return int(load().nb_batch_config_count())


def batch_config_info(cfg: int) -> dict:
    """Batch shape `cfg`: block, unroll, tile and scheme ("targets" or
    "symmetric", where tile is the super-tile length)."""
    lib = load()
    vals = [ctypes.c_int() for _ in range(3)]
    check(lib.nb_batch_config_info(cfg, *[ctypes.byref(v) for v in vals]))
    info = dict(zip(("block", "unroll", "tile"), (v.value for v in vals)))
    scheme = int(lib.nb_batch_config_scheme(cfg))
    if scheme < 0:
        check(NB_ERR_INVALID_ARG)
    info["scheme"] = {0: "targets", 1: "symmetric"}[scheme]
    return info


class Batch:
    def __init__(self, params: Sequence[NbParams], pos, vel):
        self._lib = load()
        self.params = list(params)
        self.sizes = [int(p.num_particles) for p in self.params]
        self.total = sum(self.sizes)
        arrs = [np.ascontiguousarray(a, np.float32) for a in (*pos, *vel)]
        if any(a.shape != (self.total,) for a in arrs):
            raise ValueError("arrays must hold the sum of the systems' particle counts")
        table = (NbParams * max(1, len(self.params)))(*self.params)
        self._h = ctypes.c_void_p()
        check(self._lib.nb_batch_create(len(self.params), table, *[_p(a) for a in arrs],
                                        ctypes.byref(self._h)))

    def close(self) -> None:
        if self._h:
            self._lib.nb_batch_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, steps: int) -> float:
        check(self._lib.nb_batch_step(self._h, steps))
        return self.last_ms

    def step_direct(self, steps: int) -> float:
        check(self._lib.nb_batch_step_direct(self._h, steps))
        return self.last_ms

    def step_separate(self, steps: int) -> float:
        check(self._lib.nb_batch_step_separate(self._h, steps))
        return self.last_ms

    def autotune(self, trial_steps: int = 2) -> int:
        best = ctypes.c_int()
        check(self._lib.nb_batch_autotune(self._h, trial_steps, ctypes.byref(best)))
        return best.value

    def set_config(self, cfg: int) -> None:
        check(self._lib.nb_batch_set_config(self._h, cfg))

    @property
    def config(self) -> int:
        return int(self._lib.nb_batch_config(self._h))

    @property
    def last_ms(self) -> float:
        return float(self._lib.nb_batch_last_ms(self._h))

    def _read(self, fn):
        out = [np.empty(self.total, np.float32) for _ in range(3)]
        check(fn(self._h, *[_p(a) for a in out]))
        return tuple(out)

    def positions(self):
        return self._read(self._lib.nb_batch_read_pos)

    def velocities(self):
        return self._read(self._lib.nb_batch_read_vel)


def plummer_batch_sizes(num_systems: int = 64, seed: int = 0, lo: int = 1024, hi: int = 32768,
                        step: int = 256) -> list:
    """System sizes drawn uniformly from {lo, lo+step, ..., hi} (seeded)."""
    rng = np.random.default_rng(seed)
    choices = np.arange(lo, hi + 1, step)
    return [int(v) for v in rng.choice(choices, size=num_systems)]
