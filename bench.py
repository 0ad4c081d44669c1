"""Benchmark of the MI355X all-pairs gravity step.

    python bench.py [--gpus 1] [--steps 100] [--warmup 4] [--n 262144] [--no-cpu]

Workload: the plain forward step at N = 262144 (1024 x 256) in fp32, disk
initial conditions of the reference generator (std::mt19937 seed 5489),
SimParam defaults.  One step = one launch of the pair kernel over all N
targets and N sources (N^2 interactions, 20 FLOP credited per interaction).

Prints one JSON line.  `value` is Ginteractions/s of the timed steps with
the particles resident in HBM.  `roofline` prices the pair kernel against
MI355X's fp32 vector peak (157.3 TFLOP/s) using HIP events on the launch
stream.  `cpu_baseline` times the fp32 restatement of the reference kernel
(oracle, OpenMP) on a bounded sample of the same step on the host cores.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "cuda-to-sycl-nbody_amd"))

METRIC = "Ginteractions/s fwd & fwd+adjoint at N=262144 FP32; % FP32 roofline"
FLOP_PER_INTERACTION = 20
PEAK_FP32_TFLOPS = 157.3  # MI355X published fp32 vector peak


def parse_args():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=4)
    ap.add_argument("--n", type=int, default=262144)
    ap.add_argument("--cfg", type=int, default=-1, help="forward launch shape (-1: default)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    return ap.parse_args()


def traffic_from_profile(n: int):
    """HBM bytes per launch of the pair kernel from the committed rocprofv3
    PMC summary (FETCH_SIZE + WRITE_SIZE, calibrated), if one exists for n."""
    f = ROOT / "profiles" / f"forward_n{n}_hbm_traffic.txt"
    if not f.exists():
        return None
    m = re.search(r"^bytes_per_launch:\s*([0-9.eE+]+)", f.read_text(), re.M)
    return float(m.group(1)) if m else None


def cpu_baseline(pos, vel, n: int, seconds: float) -> dict:
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as orc

    p = orc.params(num_particles=n)
    probe = 256
    t0 = time.perf_counter()
    orc.step_f32(p, pos, vel, 0, probe)
    dt = time.perf_counter() - t0
    targets = int(min(n, max(probe, probe * seconds / max(dt, 1e-6))))
    targets -= targets % 16 or 0
    t0 = time.perf_counter()
    orc.step_f32(p, pos, vel, 0, targets)
    dt = time.perf_counter() - t0
    return {
        "value": targets * n / dt / 1e9,
        "unit": "Ginteractions/s",
        "cores": orc.threads(),
        "kind": "port",
        "sample": f"one fp32 step of targets [0,{targets}) of N={n} against all {n} sources "
                  f"(reference-order restatement, OpenMP over targets, {dt:.1f} s)",
    }


def main() -> int:
    args = parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the all-pairs step "
              "is measured on a single MI355X", file=sys.stderr)
        return 2

    import numpy as np
    import torch

    import nbody_amd as nb

    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        return 1
    n = args.n
    pos, vel = nb.disk_ic(n)
    params = nb.default_params(num_particles=n, sim_iterations_per_frame=args.warmup)
    sim = nb.Simulator(params, pos, vel)
    if args.cfg >= 0:
        sim.set_config(args.cfg)
    if args.warmup > 0:
        sim.step_sim()  # warm-up frame (untimed)

    # One frame of exactly `steps` launches: nb_sim_step records HIP events
    # on its launch stream around the launches and synchronises at the end.
    sim.set_frame_iterations(args.steps)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sim.step_sim()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ev_ms = sim.last_kernel_ms
    cfg = nb.forward_config_info(sim.config)
    sim.close()

    inter = float(n) * float(n)
    value = args.steps * inter / wall / 1e9
    per_launch_ms = ev_ms / args.steps
    achieved = FLOP_PER_INTERACTION * inter / (per_launch_ms * 1e-3) / 1e12
    line = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Ginteractions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3 / args.steps, 4),
        "higher_is_better": True,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: reference disk initial conditions (std::mt19937 seed 5489)",
        "config": {
            "workload": f"plain forward step, N={n} fp32, SimParam defaults, {args.steps} steps",
            "N": n,
            "interactions_per_step": inter,
            "device": nb.device_name(),
            "launch": cfg,
        },
        "roofline": {
            "bound": "valu_fp32",
            "achieved": round(achieved, 2),
            "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(achieved / PEAK_FP32_TFLOPS, 4),
            "traffic": traffic_from_profile(n),
            "kernel_ms_per_launch": round(per_launch_ms, 4),
        },
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(pos, vel, n, args.cpu_seconds)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
