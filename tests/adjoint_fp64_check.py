"""Full-size check of the differentiable trajectory (run by hand on a GPU
host; not collected by pytest):

    python tests/adjoint_fp64_check.py [--n 16384] [--steps 200] [--every 20]

Runs nbody_amd.studies.adjoint (GPU gradient via the checkpointed adjoint,
central differences with h = 1e-3 on 32 coordinates) and compares the
gradient with the fp64 CPU adjoint of the oracle (every state stored, no
checkpoints).  Prints one JSON object.
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
    a = ap.parse_args()
    from nbody_amd import studies

    import oracle as orc

    res, arr = studies.adjoint(n=a.n, steps=a.steps, every=a.every, arrays=True)
    t0 = time.perf_counter()
    L64, gx64, gv64 = orc.traj_grad_f64(orc.params(), arr["x0"], arr["v0"], arr["target"],
                                        a.steps)
    rel = lambda u, w: float(np.linalg.norm(u - w) / np.linalg.norm(w))
    g64 = np.array([(gx64 if w == 0 else gv64)[c, i] for (w, c, i) in arr["coords"]])
    res["fp64_cpu"] = {
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
