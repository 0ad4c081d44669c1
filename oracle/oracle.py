"""ctypes loader for the CPU oracle (nbody_oracle.c).  TEST INFRASTRUCTURE:
used by tests/, __graft_entry__.smoke() and bench.py's CPU baseline only, as
the checker; the product never imports it.

Picks the AVX-512 build when the host CPU has AVX-512F, else the AVX2 build.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from ctypes import POINTER, c_double, c_int, c_int64, c_size_t, c_uint32, c_uint64, c_void_p
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
REF_DIR = HERE / "_ref"


class OrcParams(ctypes.Structure):
    _fields_ = [
        ("G", ctypes.c_float),
        ("dt", ctypes.c_float),
        ("num_particles", c_size_t),
        ("num_frames", c_size_t),
        ("sim_iterations_per_frame", c_int),
        ("damping", ctypes.c_float),
        ("dist_eps", ctypes.c_float),
        ("gw_size", c_int),
        ("calc_method", c_int),
    ]


def params(**kw) -> OrcParams:
    """SimParam defaults (src/sim_param.cpp:12-22) with overrides."""
    p = OrcParams(G=2.0, dt=0.005, num_particles=12800, num_frames=2**64 - 1,
                  sim_iterations_per_frame=4, damping=0.999998, dist_eps=1.0e-7, gw_size=64,
                  calc_method=0)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def _has_avx512() -> bool:
    try:
        with open("/proc/cpuinfo") as f:
            return " avx512f" in f.read()
    except OSError:
        return False


def build() -> None:
    subprocess.run(["make", "-s", "-C", str(HERE), "all"], check=True)


_lib = None


def load() -> ctypes.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    name = "liboracle_v4.so" if _has_avx512() else "liboracle_v3.so"
    path = HERE / name
    if not path.exists():
        build()
    lib = ctypes.CDLL(str(path))
    vp = c_void_p
    P = POINTER(OrcParams)
    lib.orc_mt_nth.restype = c_uint32
    lib.orc_mt_nth.argtypes = [c_uint32, c_uint64]
    lib.orc_disk.argtypes = [c_size_t, c_uint32, c_int] + [vp] * 6
    lib.orc_disk_gen.argtypes = [c_size_t, c_uint32] + [vp] * 6
    lib.orc_step_f32.argtypes = [P] + [vp] * 9 + [c_size_t, c_size_t, c_size_t]
    lib.orc_step_f32.restype = None
    lib.orc_forces_f64.argtypes = [vp] * 3 + [c_size_t, vp, c_size_t, c_double, c_int] + [vp] * 3
    lib.orc_forces_f64.restype = None
    lib.orc_step_f64.argtypes = [P] + [vp] * 6 + [c_size_t]
    lib.orc_step_f64.restype = None
    lib.orc_energy_momentum_f64.argtypes = [P] + [vp] * 6 + [c_size_t, vp]
    lib.orc_energy_momentum_f64.restype = None
    lib.orc_adjoint_step_f64.argtypes = [P] + [vp] * 15 + [c_size_t]
    lib.orc_adjoint_step_f64.restype = None
    lib.orc_traj_grad_f64.argtypes = [P, vp, vp, vp, c_size_t, c_int, vp, vp]
    lib.orc_traj_grad_f64.restype = c_double
    _lib = lib
    return lib


def threads() -> int:
    return int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))


def _p(a: np.ndarray) -> c_void_p:
    assert a.flags.c_contiguous
    return c_void_p(a.ctypes.data)


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32)


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def mt_nth(seed: int, n: int) -> int:
    return int(load().orc_mt_nth(seed, n))


def disk(n: int, seed: int = 5489, flags: int = 0):
    a = [np.empty(n, np.float32) for _ in range(6)]
    load().orc_disk(n, seed, flags, *[_p(v) for v in a])
    return tuple(a[:3]), tuple(a[3:])


def disk_gen(n: int, seed: int = 5489):
    a = [np.empty(n, np.float32) for _ in range(6)]
    load().orc_disk_gen(n, seed, *[_p(v) for v in a])
    return tuple(a[:3]), tuple(a[3:])


def step_f32(p: OrcParams, pos, vel, i0: int = 0, i1: int | None = None):
    """One reference-order fp32 step.  Returns (next_pos, new_vel) for
    targets [i0, i1) (other entries of the returned arrays are copies of the
    inputs)."""
    pos = [_f32(a) for a in pos]
    vel = [_f32(a).copy() for a in vel]
    n = pos[0].size
    i1 = n if i1 is None else i1
    nxt = [a.copy() for a in pos]
    load().orc_step_f32(ctypes.byref(p), *[_p(a) for a in pos], *[_p(a) for a in vel],
                        *[_p(a) for a in nxt], n, i0, i1)
    return tuple(nxt), tuple(vel)


def run_f32(p: OrcParams, pos, vel, steps: int):
    for _ in range(steps):
        pos, vel = step_f32(p, pos, vel)
    return pos, vel


def forces_f64(pos, idx=None, eps: float = 1e-7, skip_self: bool = True):
    pos = [_f32(a) for a in pos]
    n = pos[0].size
    idx = np.arange(n, dtype=np.int64) if idx is None else np.ascontiguousarray(idx, np.int64)
    out = [np.empty(idx.size, np.float64) for _ in range(3)]
    load().orc_forces_f64(*[_p(a) for a in pos], n, _p(idx), idx.size, float(eps),
                          int(skip_self), *[_p(a) for a in out])
    return tuple(out)


def step_f64(p: OrcParams, pos, vel):
    pos = [_f64(a).copy() for a in pos]
    vel = [_f64(a).copy() for a in vel]
    load().orc_step_f64(ctypes.byref(p), *[_p(a) for a in pos], *[_p(a) for a in vel],
                        pos[0].size)
    return tuple(pos), tuple(vel)


def energy_momentum_f64(p: OrcParams, pos, vel) -> np.ndarray:
    pos = [_f32(a) for a in pos]
    vel = [_f32(a) for a in vel]
    out = np.empty(5, np.float64)
    load().orc_energy_momentum_f64(ctypes.byref(p), *[_p(a) for a in pos],
                                   *[_p(a) for a in vel], pos[0].size, _p(out))
    return out


def adjoint_step_f64(p: OrcParams, pos, lx_next, lv_next):
    pos = [_f64(a) for a in pos]
    lxn = [_f64(a) for a in lx_next]
    lvn = [_f64(a) for a in lv_next]
    n = pos[0].size
    lx = [np.empty(n) for _ in range(3)]
    lv = [np.empty(n) for _ in range(3)]
    load().orc_adjoint_step_f64(ctypes.byref(p), *[_p(a) for a in pos], *[_p(a) for a in lxn],
                                *[_p(a) for a in lvn], *[_p(a) for a in lx],
                                *[_p(a) for a in lv], n)
    return tuple(lx), tuple(lv)


def traj_grad_f64(p: OrcParams, x0, v0, target, steps: int, grad: bool = True):
    """x0, v0, target: (3, n) arrays.  Returns (loss, gx, gv) with gx, gv (3, n)."""
    x0 = _f64(x0)
    v0 = _f64(v0)
    target = _f64(target)
    n = x0.shape[1]
    gx = np.empty((3, n)) if grad else None
    gv = np.empty((3, n)) if grad else None
    loss = load().orc_traj_grad_f64(ctypes.byref(p), _p(x0), _p(v0), _p(target), n, steps,
                                    _p(gx) if grad else None, _p(gv) if grad else None)
    return float(loss), gx, gv
