"""Full-size check of the differentiable trajectory (run by hand on a GPU
host; not collected by pytest):

    python tests/adjoint_fp64_check.py [--n 16384] [--steps 200] [--every 20]

Runs nbody_amd.studies.adjoint (GPU gradient via the checkpointed adjoint,
central differences with h = 1e-3 on 32 coordinates) and compares the
gradient with
  * the fp64 CPU adjoint along the GPU's own fp32 states (same trajectory,
    same residual): the adjoint's arithmetic error;
  * the fp64 CPU trajectory + adjoint from the same initial state (every state
    stored): differs also by the fp32 rounding of the trajectory, which the
    target-matching loss (residual ~1e-3 on positions ~50) amplifies.
Prints one JSON object.
"""
import argparse
import json
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT / "cuda-to-sycl-nbody_amd"))
sys.path.insert(0, str(ROOT / "oracle"))


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=16384)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--every", type=int, default=20)
    ap.add_argument("--eps", type=float, help="dist_eps (default: SimParam's)")
    ap.add_argument("--target", choices=["perturbed", "far"], default="perturbed")
    a = ap.parse_args()
    from nbody_amd import studies

    import oracle as orc

    import torch

    import nbody_amd as nb

    res, arr = studies.adjoint(n=a.n, steps=a.steps, every=a.every, arrays=True, eps=a.eps,
                               target=a.target)
    params = nb.default_params(num_particles=a.n)
    if a.eps is not None:
        params.dist_eps = a.eps
    p = orc.params() if a.eps is None else orc.params(dist_eps=a.eps)
    P = torch.from_numpy(arr["x0"]).cuda()
    V = torch.from_numpy(arr["v0"]).cuda()
    states = [arr["x0"]]
    for _ in range(a.steps):  # the forward pass's kernel and launch shape
        N = torch.empty_like(P)
        nb.forward_step(params, list(P), list(V), list(N))
        P = N
        states.append(P.cpu().numpy())
    t0 = time.perf_counter()
    lx = 2.0 * (states[-1].astype(np.float64) - arr["target"]) / a.n
    lv = np.zeros_like(lx)
    for t in range(a.steps - 1, -1, -1):
        lx, lv = orc.adjoint_step_f64(p, states[t], lx, lv)
        lx, lv = np.stack(lx), np.stack(lv)
    rel = lambda u, w: float(np.linalg.norm(u - w) / np.linalg.norm(w))
    res["fp64_same_states"] = {"seconds": time.perf_counter() - t0,
                               "grad_x_rel_err": rel(arr["gx"], lx),
                               "grad_v_rel_err": rel(arr["gv"], lv)}
    t0 = time.perf_counter()
    L64, gx64, gv64 = orc.traj_grad_f64(p, arr["x0"], arr["v0"], arr["target"], a.steps)
    g64 = np.array([(gx64 if w == 0 else gv64)[c, i] for (w, c, i) in arr["coords"]])
    res["fp64_independent_trajectory"] = {
        "seconds": time.perf_counter() - t0,
        "threads": orc.threads(),
        "loss": L64,
        "grad_x_rel_err": rel(arr["gx"], gx64),
        "grad_v_rel_err": rel(arr["gv"], gv64),
        "fd_vs_fp64_grad_rel_err": rel(arr["fd"], g64),
    }
    print(json.dumps(res), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
