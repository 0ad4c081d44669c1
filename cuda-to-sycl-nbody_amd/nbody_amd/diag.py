"""fp64 diagnostics on the GPU: oracle forces on a sampled subset and the
total energy / momentum (libnbody_amd nb_forces_f64_dev,
nb_energy_momentum_dev)."""
from __future__ import annotations

import ctypes

from ._lib import NbParams, check, load
from .adjoint import _stream, _vec3


def forces_f64(params: NbParams, pos, idx, stream=None):
    """fp64 force on targets idx (int32 CUDA tensor) from all sources in pos
    ((3, n) float32 CUDA tensor).  Returns a (3, m) float64 CUDA tensor."""
    import torch

    n = int(pos[0].numel())
    if not idx.is_cuda:
        raise ValueError("idx must be a CUDA tensor")
    idx = idx.to(torch.int32).contiguous()
    m = int(idx.numel())
    out = torch.empty(3, m, dtype=torch.float64, device=pos.device)
    check(load().nb_forces_f64_dev(ctypes.byref(params), ctypes.byref(_vec3(pos, n)),
                                   ctypes.c_void_p(idx.data_ptr()), m,
                                   *[ctypes.c_void_p(out[c].data_ptr()) for c in range(3)], n,
                                   _stream(stream)))
    return out


def energy_momentum(params: NbParams, pos, vel, stream=None) -> dict:
    """fp64 kinetic, potential, total energy and momentum of the state."""
    import torch

    n = int(pos[0].numel())
    out = torch.empty(5, dtype=torch.float64, device=pos.device)
    check(load().nb_energy_momentum_dev(ctypes.byref(params), ctypes.byref(_vec3(pos, n)),
                                        ctypes.byref(_vec3(vel, n)), n,
                                        ctypes.c_void_p(out.data_ptr()), _stream(stream)))
    k, w, px, py, pz = out.cpu().tolist()
    return {"kinetic": k, "potential": w, "total": k + w, "momentum": (px, py, pz)}
