"""Python access to the MI355X all-pairs gravity step (libnbody_amd).

    import nbody_amd as nb
    pos, vel = nb.disk_ic(25600)                # reference disk, host arrays
    sim = nb.Simulator(nb.default_params(num_particles=25600), pos, vel)
    sim.step_sim()                              # one frame, like stepSim()
    x, y, z = sim.positions()

Device-level launches take PyTorch tensors as device memory:
    nb.forward_step(params, pos, vel, nxt)      # pos/vel/nxt: 3 float32 CUDA tensors each

All work runs in the HIP library; nothing here computes the physics.
"""
from __future__ import annotations

import ctypes
from typing import Sequence

import numpy as np

from .diag import energy_momentum, forces_f64
from .batch import Batch, batch_config_count, batch_config_info, plummer_batch_sizes
from .adjoint import (adjoint_config_count, adjoint_config_info, adjoint_default_config,
                      adjoint_step, fwd_adjoint_step, trajectory)
from ._lib import (DEFAULT_SEED, DRIVER_PATH, LIB_DIR, LIB_PATH, NB_ACC_KAHAN, NB_ACC_PLAIN,
                   NB_CALC_BRANCH, NB_CALC_PREDICATED, NB_ERR_INVALID_ARG, NB_ERR_NO_DEVICE,
                   NB_IC_LENGTH_F32, NB_IC_TRIG_DOUBLE, NB_OK, SIM_LIB_PATH, NbError, NbParams,
                   NbVec3, check, default_params, load)

__all__ = [
    "DEFAULT_SEED", "DRIVER_PATH", "LIB_DIR", "LIB_PATH", "SIM_LIB_PATH", "NB_ACC_KAHAN",
    "NB_ACC_PLAIN", "NB_CALC_BRANCH", "NB_CALC_PREDICATED", "NB_ERR_INVALID_ARG",
    "NB_ERR_NO_DEVICE", "NB_IC_LENGTH_F32", "NB_IC_TRIG_DOUBLE", "NB_OK", "NbError", "NbParams",
    "Simulator", "check", "default_params", "device_count", "device_name", "disk_ic",
    "disk_gen_ic", "forward_config_count", "forward_config_info", "forward_step", "load",
    "plummer_ic", "adjoint_config_count", "adjoint_config_info", "adjoint_step",
    "fwd_adjoint_step", "trajectory", "NbVec3", "Batch", "batch_config_count",
    "batch_config_info", "plummer_batch_sizes", "energy_momentum", "forces_f64", "pack_vec4",
    "scratch_release", "scratch_held", "adjoint_default_config", "forward_default_config",
]


def forward_default_config(n: int, params: NbParams | None = None,
                           acc: int = NB_ACC_PLAIN) -> int:
    """The forward shape the library picks for n particles (SimParam
    defaults unless params is given)."""
    p = params if params is not None else default_params(num_particles=n)
    cfg = int(load().nb_forward_default_config(ctypes.byref(p), n, acc))
    if cfg < 0:
        check(NB_ERR_INVALID_ARG)
    return cfg


def scratch_release() -> None:
    """Frees the library's per-stream scratch buffers on the current device
    (after waiting for all work on it)."""
    check(load().nb_scratch_release())


def scratch_held() -> int:
    """Bytes of per-stream scratch the library holds on the current device."""
    b = ctypes.c_size_t()
    check(load().nb_scratch_held(ctypes.byref(b)))
    return int(b.value)


def _ptr(a: np.ndarray) -> ctypes.c_void_p:
    if a.dtype != np.float32 or not a.flags.c_contiguous:
        raise ValueError("expected a contiguous float32 array")
    return ctypes.c_void_p(a.ctypes.data)


def _six(n: int) -> list[np.ndarray]:
    return [np.empty(n, dtype=np.float32) for _ in range(6)]


def disk_ic(n: int, seed: int = DEFAULT_SEED, flags: int = 0):
    """Disk initial conditions of DiskGalaxySimulator (src/simulator.cu:131-158).
    Returns ((x, y, z), (vx, vy, vz)) float32 arrays."""
    a = _six(n)
    check(load().nb_ic_disk(n, seed, flags, *[_ptr(v) for v in a]))
    return tuple(a[:3]), tuple(a[3:])


def disk_gen_ic(n: int, seed: int = DEFAULT_SEED):
    """Disk in gen.cpp's draw order (src/gen.cpp:15-34)."""
    a = _six(n)
    check(load().nb_ic_disk_gen(n, seed, *[_ptr(v) for v in a]))
    return tuple(a[:3]), tuple(a[3:])


def plummer_ic(n: int, seed: int, a: float = 1.0, g_mass: float = 1.0):
    """Plummer sphere (scale radius a, virial velocities for G*M = g_mass)."""
    arr = _six(n)
    check(load().nb_ic_plummer(n, seed, a, g_mass, *[_ptr(v) for v in arr]))
    return tuple(arr[:3]), tuple(arr[3:])


def device_count() -> int:
    return int(load().nb_device_count())


def device_name() -> str:
    buf = ctypes.create_string_buffer(256)
    check(load().nb_device_name(buf, 256))
    return buf.value.decode()


def forward_config_count() -> int:
    return int(load().nb_forward_config_count())


SCHEMES = {0: "targets", 1: "symmetric", 2: "symmetric_tiles"}


def forward_config_info(cfg: int) -> dict:
    """Launch shape `cfg`: block, unroll, tile, split and scheme ("targets":
    each thread sums all sources for its targets; "symmetric": each pair term
    evaluated once and applied to both particles, where tile is the
    super-tile length and split the super-tiles per work item;
    "symmetric_tiles": the same with one tile pair per wave, tile = the tile
    length)."""
    lib = load()
    vals = [ctypes.c_int() for _ in range(4)]
    check(lib.nb_forward_config_info(cfg, *[ctypes.byref(v) for v in vals]))
    info = dict(zip(("block", "unroll", "tile", "split"), (v.value for v in vals)))
    scheme = int(lib.nb_forward_config_scheme(cfg))
    if scheme < 0:
        check(NB_ERR_INVALID_ARG)
    info["scheme"] = SCHEMES[scheme]
    return info


class Simulator:
    """The owning simulator handle (nb_sim_*): DiskGalaxySimulator's contract."""

    def __init__(self, params: NbParams, pos: Sequence[np.ndarray], vel: Sequence[np.ndarray]):
        n = params.num_particles
        arrs = [np.ascontiguousarray(a, dtype=np.float32) for a in (*pos, *vel)]
        if any(a.shape != (n,) for a in arrs):
            raise ValueError("each array must hold params.num_particles values")
        self._lib = load()
        self._h = ctypes.c_void_p()
        # a copy: set_frame_iterations updates it, not the caller's struct
        self.params = NbParams.from_buffer_copy(params)
        check(self._lib.nb_sim_create(ctypes.byref(self.params), *[_ptr(a) for a in arrs],
                                      ctypes.byref(self._h)))

    def close(self) -> None:
        if self._h:
            self._lib.nb_sim_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step_sim(self) -> None:
        check(self._lib.nb_sim_step(self._h))

    def advance(self, iters: int) -> None:
        check(self._lib.nb_sim_advance(self._h, iters))

    def sync(self) -> None:
        check(self._lib.nb_sim_sync(self._h))

    def set_accumulation(self, mode: int) -> None:
        check(self._lib.nb_sim_set_accumulation(self._h, mode))

    def set_frame_iterations(self, iters: int) -> None:
        check(self._lib.nb_sim_set_frame_iterations(self._h, iters))
        self.params.sim_iterations_per_frame = iters

    def set_config(self, cfg: int) -> None:
        check(self._lib.nb_sim_set_config(self._h, cfg))

    def set_graph(self, enable: bool) -> None:
        """Replay frames of two or more steps from a captured hipGraph (the
        default) or launch every step directly."""
        check(self._lib.nb_sim_set_graph(self._h, int(bool(enable))))

    @property
    def config(self) -> int:
        return int(self._lib.nb_sim_config(self._h))

    def _read(self, fn):
        n = self.params.num_particles
        out = [np.empty(n, dtype=np.float32) for _ in range(3)]
        check(fn(self._h, *[_ptr(a) for a in out]))
        return tuple(out)

    def positions(self):
        return self._read(self._lib.nb_sim_read_pos)

    def staged(self):
        """Positions and velocities through the pinned staging buffer:
        ((x, y, z), (vx, vy, vz)) copies of the six arrays."""
        check(self._lib.nb_sim_stage_readback(self._h))
        ptrs = (ctypes.c_void_p * 6)()
        check(self._lib.nb_sim_staged(self._h, ptrs))
        n = self.params.num_particles
        out = []
        for p in ptrs:
            a = np.empty(n, np.float32)
            if n:
                ctypes.memmove(a.ctypes.data, p, 4 * n)
            out.append(a)
        return tuple(out[:3]), tuple(out[3:])

    def read_vec4(self, which: int = 0, w: float = 1.0) -> np.ndarray:
        """(n, 4) float32: positions (which=0) or velocities (1) as vec4."""
        out = np.empty((self.params.num_particles, 4), np.float32)
        check(self._lib.nb_sim_read_vec4(self._h, which, w, ctypes.c_void_p(out.ctypes.data)))
        return out

    def velocities(self):
        return self._read(self._lib.nb_sim_read_vel)

    @property
    def last_step_ms(self) -> float:
        return float(self._lib.nb_sim_last_step_ms(self._h))

    @property
    def last_kernel_ms(self) -> float:
        return float(self._lib.nb_sim_last_kernel_ms(self._h))

    @property
    def num_particles(self) -> int:
        return int(self._lib.nb_sim_num_particles(self._h))


def _dev_ptrs(ts, n: int, name: str) -> list[ctypes.c_void_p]:
    if len(ts) != 3:
        raise ValueError(f"{name}: expected 3 tensors (x, y, z)")
    out = []
    for t in ts:
        if not t.is_cuda or str(t.dtype) != "torch.float32" or not t.is_contiguous():
            raise ValueError(f"{name}: tensors must be contiguous float32 on the GPU")
        if t.numel() < n:
            raise ValueError(f"{name}: tensor holds {t.numel()} < {n} values")
        out.append(ctypes.c_void_p(t.data_ptr()))
    return out


def pack_vec4(soa, w: float = 1.0, stream=None):
    """(3, n) float32 CUDA tensor -> (n, 4) float32 CUDA tensor of
    vec4(x, y, z, w), packed by the HIP kernel."""
    import torch

    from .adjoint import _stream, _vec3

    n = int(soa[0].numel())
    out = torch.empty(n, 4, dtype=torch.float32, device=soa.device)
    check(load().nb_pack_vec4_dev(ctypes.byref(_vec3(soa, n)), w, ctypes.c_void_p(out.data_ptr()),
                                  n, _stream(stream)))
    return out


def forward_step(params: NbParams, pos, vel, nxt, n: int | None = None, cfg: int = -1,
                 acc: int = NB_ACC_PLAIN, stream=None) -> None:
    """One stream-ordered step on device tensors: reads pos, updates vel in
    place, writes nxt (3 tensors each).  stream: a torch.cuda.Stream or None
    (the current stream)."""
    import torch

    n = int(pos[0].numel()) if n is None else int(n)
    ptrs = _dev_ptrs(pos, n, "pos") + _dev_ptrs(vel, n, "vel") + _dev_ptrs(nxt, n, "nxt")
    s = stream if stream is not None else torch.cuda.current_stream()
    check(load().nb_forward_step_dev(ctypes.byref(params), *ptrs, n, cfg, acc,
                                     ctypes.c_void_p(s.cuda_stream)))
