"""Adjoint of the all-pairs step and the differentiable trajectory.

    x_T, v_T = nbody_amd.trajectory(x0, v0, params, steps=200, every=20)
    loss = ((x_T - target) ** 2).sum(0).mean()
    loss.backward()            # x0.grad, v0.grad from the adjoint kernels

x0, v0 are (3, n) float32 CUDA tensors (rows x, y, z).  The forward pass runs
the plain forward kernel and keeps the state every `every` steps; backward
recomputes each segment's positions from its checkpoint and runs the
hand-written adjoint kernel in reverse (libnbody_amd, nb_traj_*).
"""
from __future__ import annotations

import ctypes

from ._lib import NB_ERR_INVALID_ARG, NbParams, NbVec3, check, load


def _vec3(t, n: int) -> NbVec3:
    """A (3, n) contiguous float32 CUDA tensor, or a sequence of 3 tensors,
    each row holding at least n values."""
    if len(t) != 3:
        raise ValueError("expected 3 rows (x, y, z)")
    rows = [t[0], t[1], t[2]]
    for r in rows:
        if not r.is_cuda or str(r.dtype) != "torch.float32" or not r.is_contiguous():
            raise ValueError("expected contiguous float32 CUDA tensors")
        if r.numel() < n:
            raise ValueError(f"a row holds {r.numel()} < {n} values")
    return NbVec3(*[ctypes.c_void_p(r.data_ptr()) for r in rows])


def _stream(stream):
    import torch

    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def adjoint_step(params: NbParams, pos, lam_x_next, lam_v_next, lam_x, lam_v, cfg: int = -1,
                 stream=None) -> None:
    """Cotangents of (x_n, v_n) from those of (x_{n+1}, v_{n+1}) at positions pos."""
    n = int(pos[0].numel())
    vs = [_vec3(t, n) for t in (pos, lam_x_next, lam_v_next, lam_x, lam_v)]
    check(load().nb_adjoint_step_dev(ctypes.byref(params), *[ctypes.byref(v) for v in vs], n,
                                     cfg, _stream(stream)))


def fwd_adjoint_step(params: NbParams, pos, vel, pos_next, lam_x_next, lam_v_next, lam_x, lam_v,
                     cfg: int = -1, stream=None) -> None:
    """Forward step (vel in place, pos_next written) and adjoint step at pos, one launch."""
    n = int(pos[0].numel())
    vs = [_vec3(t, n) for t in (pos, vel, pos_next, lam_x_next, lam_v_next, lam_x, lam_v)]
    check(load().nb_fwd_adjoint_step_dev(ctypes.byref(params), *[ctypes.byref(v) for v in vs],
                                         n, cfg, _stream(stream)))


def adjoint_default_config(n: int, params: NbParams | None = None) -> int:
    """The adjoint shape the library picks for n particles (SimParam
    defaults unless params is given)."""
    from ._lib import default_params

    p = params if params is not None else default_params(num_particles=n)
    cfg = int(load().nb_adjoint_default_config(ctypes.byref(p), n))
    if cfg < 0:
        check(NB_ERR_INVALID_ARG)
    return cfg


def adjoint_config_count() -> int:
    return int(load().nb_adjoint_config_count())


def adjoint_config_info(cfg: int) -> dict:
    """Adjoint launch shape `cfg` (fields as forward_config_info's)."""
    lib = load()
    vals = [ctypes.c_int() for _ in range(4)]
    check(lib.nb_adjoint_config_info(cfg, *[ctypes.byref(v) for v in vals]))
    info = dict(zip(("block", "unroll", "tile", "split"), (v.value for v in vals)))
    scheme = int(lib.nb_adjoint_config_scheme(cfg))
    if scheme < 0:
        check(NB_ERR_INVALID_ARG)
    info["scheme"] = {0: "targets", 1: "symmetric", 2: "symmetric_tiles"}[scheme]
    return info


def _ptr(t) -> ctypes.c_void_p:
    return ctypes.c_void_p(t.data_ptr())


def _make_fn():
    import torch

    class TrajectoryFn(torch.autograd.Function):
        @staticmethod
        def forward(ctx, x0, v0, params, steps, every):
            lib = load()
            n = x0.shape[1]
            dev = x0.device
            state0 = torch.cat([x0.detach(), v0.detach()]).to(torch.float32).contiguous()
            ck = torch.empty(max(1, lib.nb_traj_checkpoint_floats(n, steps, every)),
                             dtype=torch.float32, device=dev)
            ws = torch.empty(max(1, lib.nb_traj_workspace_floats(n, every)),
                             dtype=torch.float32, device=dev)
            out = torch.empty(6, n, dtype=torch.float32, device=dev)
            check(lib.nb_traj_forward_dev(ctypes.byref(params), _ptr(state0), _ptr(out), steps,
                                          every, _ptr(ck), _ptr(ws), n, _stream(None)))
            ctx.save_for_backward(ck)
            ctx.params = NbParams.from_buffer_copy(params)
            ctx.steps, ctx.every, ctx.n = steps, every, n
            return out[:3], out[3:]

        @staticmethod
        def backward(ctx, gx, gv):
            lib = load()
            (ck,) = ctx.saved_tensors
            n, dev = ctx.n, ck.device
            z = torch.zeros(3, n, dtype=torch.float32, device=dev)
            lam_T = torch.cat([z if gx is None else gx.to(torch.float32),
                               z if gv is None else gv.to(torch.float32)]).contiguous()
            ws = torch.empty(max(1, lib.nb_traj_workspace_floats(n, ctx.every)),
                             dtype=torch.float32, device=dev)
            lam0 = torch.empty(6, n, dtype=torch.float32, device=dev)
            check(lib.nb_traj_backward_dev(ctypes.byref(ctx.params), _ptr(ck), _ptr(lam_T),
                                           _ptr(lam0), ctx.steps, ctx.every, _ptr(ws), n,
                                           _stream(None)))
            return lam0[:3], lam0[3:], None, None, None

    return TrajectoryFn


_FN = None


def trajectory(x0, v0, params: NbParams, steps: int, every: int = 20):
    """Differentiable T-step trajectory: returns (x_T, v_T), each (3, n)."""
    global _FN
    for t in (x0, v0):
        if not t.is_cuda or t.dim() != 2 or t.shape[0] != 3:
            raise ValueError("x0 and v0 must be (3, n) CUDA tensors")
    if x0.shape != v0.shape or x0.device != v0.device:
        raise ValueError("x0 and v0 must have the same shape and device")
    if _FN is None:
        _FN = _make_fn()
    return _FN.apply(x0, v0, params, int(steps), int(every))
