"""Measurement studies beside the headline benchmark.

    python -m nbody_amd.studies precision [--n 1048576] [--steps 100]
    python -m nbody_amd.studies adjoint   [--n 16384] [--steps 200] [--every 20]
    python -m nbody_amd.studies shapes    [--sizes 16384 65536 262144]

precision: N particles, half a Plummer sphere (scale radius 1, virial
  velocities for the simulation's own G and unit masses, centred on the disk
  mid-plane) and half the reference disk, seed 7.  For plain fp32 (the
  library's symmetric scheme and the target-sharded one) and
  Kahan-compensated fp32 force accumulation: force error percentiles against
  the fp64 GPU oracle on a random 4096-particle sample, throughput over
  `steps` steps, and the fp64 total energy and momentum before and after.
adjoint: the reference disk, a target made by a forward run from initial
  positions perturbed by Gaussian sigma = 1e-3, loss = mean squared distance
  of the final positions; gradient w.r.t. initial positions and velocities
  from the checkpointed adjoint, compared with central differences
  (h = 1e-3, fp32 GPU forward runs) on 32 coordinates.  (The comparison with
  an fp64 CPU adjoint at this workload is a GPU test,
  tests/test_gpu_workloads.py.)
shapes: every compiled launch shape timed at each size (forward step plain
  and Kahan, adjoint plain and fused), fastest first: the measurements the
  library's default shapes come from.

Prints one JSON object.
"""
from __future__ import annotations

import argparse
import json
import sys
import time
import numpy as np

from . import _lib
from ._lib import NbParams


def _torch():
    import torch

    if not torch.cuda.is_available():
        raise SystemExit("the studies need a GPU")
    return torch


def mixed_ics(n: int, seed: int = 7, G: float = 2.0, a: float = 1.0):
    """Half Plummer sphere (scale radius a, virial for G * n/2), half
    reference disk."""
    from . import disk_ic, plummer_ic

    half = n // 2
    (px, py, pz), (vx, vy, vz) = plummer_ic(half, seed=seed, a=a, g_mass=G * half)
    (dx, dy, dz), (ux, uy, uz) = disk_ic(n - half, seed=seed)
    pz = pz + np.float32(2.0)  # the disk spans z in [0, 4)
    cat = lambda a, b: np.ascontiguousarray(np.concatenate([a, b]), np.float32)
    return ((cat(px, dx), cat(py, dy), cat(pz, dz)), (cat(vx, ux), cat(vy, uy), cat(vz, uz)))


def _percentiles(e: np.ndarray) -> dict:
    return {f"p{q}": float(np.percentile(e, q)) for q in (50, 90, 99, 99.9)} | {
        "max": float(e.max())}


def precision(n: int = 1048576, steps: int = 100, seed: int = 7, sample: int = 4096,
              plummer_a: float = 1.0, damping: float | None = None,
              eps: float | None = None) -> dict:
    import nbody_amd as nb

    torch = _torch()
    dev = torch.device("cuda:0")
    pos, vel = mixed_ics(n, seed, a=plummer_a)
    params = nb.default_params(num_particles=n)
    if damping is not None:
        params.damping = damping
    if eps is not None:
        params.dist_eps = eps
    P0 = torch.from_numpy(np.stack(pos)).to(dev)
    V0 = torch.from_numpy(np.stack(vel)).to(dev)
    rng = np.random.default_rng(seed)
    idx = torch.from_numpy(np.sort(rng.choice(n, size=sample, replace=False)).astype(np.int32)).to(dev)

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    f64 = nb.forces_f64(params, P0, idx)
    torch.cuda.synchronize()
    t_f64 = time.perf_counter() - t0
    f64 = f64.cpu().numpy()
    norm64 = np.linalg.norm(f64, axis=0)

    # fp32 force of a mode: one step with G = dt = 1, damping = 0, v = 0 gives v' = F
    unit = nb.default_params(num_particles=n, G=1.0, dt=1.0, damping=0.0,
                             dist_eps=params.dist_eps)
    out = {"n": n, "steps": steps, "seed": seed, "sample": sample,
           "damping": float(params.damping), "dist_eps": float(params.dist_eps),
           "ics": f"half Plummer (a={plummer_a:g}, virial for G*N/2, centred at z=2) "
                  "+ half reference disk",
           "fp64_oracle": {"seconds": t_f64,
                           "value": sample * n / t_f64 / 1e9, "unit": "Ginteractions/s"}}
    e0 = nb.energy_momentum(params, P0, V0)
    out["energy_initial"] = e0
    targets = next(c for c in range(nb.forward_config_count())
                   if nb.forward_config_info(c) == {"block": 256, "unroll": 4, "tile": 1024,
                                                    "split": 8, "scheme": "targets"})
    modes = (("fp32", nb.NB_ACC_PLAIN, -1),            # library default (symmetric)
             ("fp32_targets", nb.NB_ACC_PLAIN, targets),
             ("fp32_kahan", nb.NB_ACC_KAHAN, -1),      # library default (symmetric)
             ("fp32_kahan_targets", nb.NB_ACC_KAHAN, targets))
    for name, acc, cfg in modes:
        print(f"precision: {name} ...", file=sys.stderr, flush=True)
        V = torch.zeros_like(P0)
        N = torch.empty_like(P0)
        nb.forward_step(unit, list(P0), list(V), list(N), acc=acc, cfg=cfg)
        torch.cuda.synchronize()
        F = V[:, idx.long()].double().cpu().numpy()
        rel = np.linalg.norm(F - f64, axis=0) / norm64
        # trajectory: `steps` steps from the same state, timed
        sim = nb.Simulator(nb.default_params(num_particles=n, sim_iterations_per_frame=steps,
                                             damping=params.damping,
                                             dist_eps=params.dist_eps), pos, vel)
        sim.set_accumulation(acc)
        sim.set_config(cfg)
        launch = nb.forward_config_info(sim.config)
        sim.step_sim()
        ms = sim.last_kernel_ms / steps
        xp = torch.from_numpy(np.stack(sim.positions())).to(dev)
        xv = torch.from_numpy(np.stack(sim.velocities())).to(dev)
        sim.close()
        e1 = nb.energy_momentum(params, xp, xv)
        damp = float(np.float32(params.damping)) ** steps
        pscale = float(np.abs(np.stack(vel).astype(np.float64)).sum())
        dp = np.array(e1["momentum"]) - damp * np.array(e0["momentum"])
        print(f"precision: {name} {ms:.3f} ms/step", file=sys.stderr, flush=True)
        out[name] = {
            "launch": launch,
            "force_rel_err_vs_fp64": _percentiles(rel),
            "ms_per_step": ms,
            "value": n * float(n) / (ms * 1e-3) / 1e9,
            "unit": "Ginteractions/s",
            "energy_final": e1,
            "energy_rel_change": (e1["total"] - e0["total"]) / abs(e0["total"]),
            "momentum_drift_rel": float(np.linalg.norm(dp) / pscale),
        }
    return out


def adjoint(n: int = 16384, steps: int = 200, every: int = 20, fd_coords: int = 32,
            h: float = 1e-3, arrays: bool = False, eps: float | None = None,
            target: str = "perturbed"):
    """Returns the result dict; with arrays=True also a dict of the inputs,
    target and gradients (numpy) for an external check.

    target "perturbed": the final positions of a run from initial positions
    plus Gaussian noise (sigma 1e-3); "far": the initial positions shifted by
    10 units in every coordinate (a residual far above fp32 rounding, for a
    well-conditioned finite-difference check).  eps: dist_eps (None: the
    SimParam default)."""
    import nbody_amd as nb

    torch = _torch()
    dev = torch.device("cuda:0")
    pos, vel = nb.disk_ic(n)
    params = nb.default_params(num_particles=n)
    if eps is not None:
        params.dist_eps = eps
    x0 = torch.from_numpy(np.stack(pos)).to(dev)
    v0 = torch.from_numpy(np.stack(vel)).to(dev)
    rng = np.random.default_rng(0)
    if target == "far":
        tgt = (x0 + 10.0).contiguous()
    else:
        pert = torch.from_numpy(rng.normal(scale=1e-3, size=(3, n)).astype(np.float32)).to(dev)
        with torch.no_grad():
            tgt, _ = nb.trajectory(x0 + pert, v0, params, steps, every)
    target64 = tgt.double()

    def loss_of(x, v):
        xT, _ = nb.trajectory(x, v, params, steps, every)
        return ((xT.double() - target64) ** 2).sum(0).mean()

    x = x0.clone().requires_grad_(True)
    v = v0.clone().requires_grad_(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    L = loss_of(x, v)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    L.backward()
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    gx = x.grad.double().cpu().numpy()
    gv = v.grad.double().cpu().numpy()
    out = {"n": n, "steps": steps, "every": every, "dist_eps": float(params.dist_eps),
           "target": target, "loss": float(L.item()),
           "forward_seconds": t1 - t0, "backward_seconds": t2 - t1}

    # central differences with the fp32 GPU forward (fp64 loss)
    fd, an, coords = [], [], []
    with torch.no_grad():
        for k in range(fd_coords):
            c, i = int(rng.integers(3)), int(rng.integers(n))
            which = k % 2
            coords.append((which, c, i))
            base = (x0, v0)
            plus = [t.clone() for t in base]
            minus = [t.clone() for t in base]
            plus[which][c, i] += h
            minus[which][c, i] -= h
            hp = float(plus[which][c, i] - base[which][c, i])
            hm = float(base[which][c, i] - minus[which][c, i])
            fd.append(float(loss_of(*plus) - loss_of(*minus)) / (hp + hm))
            an.append(gx[c, i] if which == 0 else gv[c, i])
    fd, an = np.array(fd), np.array(an)
    out["central_differences"] = {
        "coords": fd_coords, "h": h,
        "rel_err_normwise": float(np.linalg.norm(an - fd) / np.linalg.norm(fd)),
        "note": "fd from fp32 GPU trajectories; fp32 rounding of x_T limits its accuracy"}
    if arrays:
        return out, {"x0": x0.cpu().numpy(), "v0": v0.cpu().numpy(),
                     "target": target64.cpu().numpy(), "gx": gx, "gv": gv,
                     "fd": fd, "coords": coords}
    return out


def shapes(sizes, kinds=("forward", "adjoint", "kahan"), cfgs=None) -> dict:
    """Times every compiled launch shape (or those listed in cfgs) on the
    reference disk at each size: the forward step (plain and Kahan
    accumulation) and the adjoint (plain and fused with the forward step).
    Shapes that refuse a size or mode are skipped.  Returns
    {kind: {n: [(cfg, info, ms), ...]}}, fastest first."""
    import nbody_amd as nb

    torch = _torch()
    dev = torch.device("cuda:0")

    def timed(fn, reps):
        fn()
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    out = {k: {} for k in kinds}
    for n in sizes:
        pos, vel = nb.disk_ic(n)
        params = nb.default_params(num_particles=n)
        P = [torch.from_numpy(a).to(dev) for a in pos]
        V = [torch.from_numpy(a).to(dev) for a in vel]
        N = [torch.empty_like(t) for t in P]
        reps = max(2, min(20, int(1e12 / (float(n) * n))))
        for kind in kinds:
            rows = []
            if kind in ("forward", "kahan"):
                acc = nb.NB_ACC_KAHAN if kind == "kahan" else nb.NB_ACC_PLAIN
                for cfg in range(nb.forward_config_count()):
                    if cfgs is not None and cfg not in cfgs:
                        continue
                    try:
                        ms = timed(lambda: nb.forward_step(params, P, V, N, cfg=cfg, acc=acc),
                                   reps)
                    except nb.NbError:
                        continue
                    rows.append((cfg, nb.forward_config_info(cfg), ms))
            else:
                rng = np.random.default_rng(0)
                lxn = torch.from_numpy(rng.normal(size=(3, n)).astype(np.float32)).to(dev)
                lvn = torch.from_numpy(rng.normal(size=(3, n)).astype(np.float32)).to(dev)
                lx, lv = torch.empty_like(lxn), torch.empty_like(lxn)
                P3 = torch.stack(P)
                V3 = torch.stack(V)
                N3 = torch.empty_like(P3)
                for cfg in range(nb.adjoint_config_count()):
                    if cfgs is not None and cfg not in cfgs:
                        continue
                    try:
                        ms = timed(lambda: nb.adjoint_step(params, P3, lxn, lvn, lx, lv,
                                                           cfg=cfg), reps)
                        fused = timed(lambda: nb.fwd_adjoint_step(params, P3, V3, N3, lxn, lvn,
                                                                  lx, lv, cfg=cfg), reps)
                    except nb.NbError:
                        continue
                    rows.append((cfg, nb.adjoint_config_info(cfg), ms, fused))
            rows.sort(key=lambda r: r[2])
            out[kind][n] = rows
            best = rows[0] if rows else None
            print(f"shapes: {kind} N={n}: {len(rows)} shapes, fastest cfg "
                  f"{best[0] if best else None} {best[2] if best else 0:.4f} ms",
                  file=sys.stderr, flush=True)
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("study", choices=["precision", "adjoint", "shapes"])
    ap.add_argument("--sizes", type=int, nargs="+", default=[16384, 65536, 262144],
                    help="shapes: particle counts")
    ap.add_argument("--kinds", nargs="+", choices=["forward", "adjoint", "kahan"],
                    default=["forward", "adjoint", "kahan"], help="shapes: what to time")
    ap.add_argument("--cfgs", type=int, nargs="+", help="shapes: only these shape indices")
    ap.add_argument("--n", type=int)
    ap.add_argument("--steps", type=int)
    ap.add_argument("--every", type=int, default=20)
    ap.add_argument("--eps", type=float, help="dist_eps (default: SimParam's)")
    ap.add_argument("--target", choices=["perturbed", "far"], default="perturbed")
    ap.add_argument("--plummer-a", type=float, default=1.0,
                    help="precision: Plummer scale radius")
    ap.add_argument("--damping", type=float, help="precision: damping (default: SimParam's)")
    a = ap.parse_args(argv)
    _lib.load()
    if a.study == "shapes":
        res = shapes(a.sizes, tuple(a.kinds), a.cfgs)
    elif a.study == "precision":
        res = precision(n=a.n or 1048576, steps=a.steps or 100, plummer_a=a.plummer_a,
                        damping=a.damping, eps=a.eps)
    else:
        res = adjoint(n=a.n or 16384, steps=a.steps or 200, every=a.every, eps=a.eps,
                      target=a.target)
    print(json.dumps(res), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
