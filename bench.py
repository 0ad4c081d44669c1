"""Benchmark of the MI355X all-pairs gravity step.

    python bench.py [--gpus 1] [--steps 100] [--warmup 4] [--n 262144] [--no-cpu]

Workload: the plain forward step at N = 262144 (1024 x 256) in fp32, disk
initial conditions of the reference generator (std::mt19937 seed 5489),
SimParam defaults.  One step = the library's launch shape for N (the
symmetric scheme: each pair term evaluated once for both particles, then an
ordered finalize that integrates) over all N targets and N sources: N^2
interactions, 20 FLOP credited per interaction (--cfg picks another shape).

Prints one JSON line.  `value` is Ginteractions/s of the timed steps with
the particles resident in HBM.  `roofline` prices the step's kernels
against MI355X's fp32 peak (157.3 TFLOP/s) using HIP events on the launch
stream; `executed_frac` prices the FLOP the hardware executes instead of the
credited ones; `traffic` is the pair kernel's HBM bytes per launch from the
committed rocprofv3 counter summary.  `cpu_baseline` times the fp32
restatement of the reference kernel (oracle, OpenMP) on a bounded sample of
the same step on the host cores.

Also in the line (same run, same N):
  fwd_adjoint  the fused forward+adjoint step (one launch, N^2 pairs) and the
               separate adjoint step, as Ginteractions/s and as a multiple of
               the forward step time;
  trajectory   a checkpointed differentiable trajectory (forward pass +
               backward pass with segment replay) per step;
  batched      64 independent Plummer systems (N_s uniform on
               {1024, 1280, ..., 32768}): the autotuned persistent kernel
               replayed from a hipGraph vs one plain launch per system.
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
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the fwd+adjoint, trajectory and batched sections")
    ap.add_argument("--batch-steps", type=int, default=200)
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB five times (the known-size "
                         "reference for rocprofv3 FETCH_SIZE / WRITE_SIZE) and exit")
    return ap.parse_args()


def calibrate_copy(mib: int) -> int:
    import torch

    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"calibrate_copy_bytes": n * 4, "launches": 5}), flush=True)
    return 0


def _events(torch, fn, reps: int) -> float:
    """Mean ms per call of fn() on the current stream (HIP events)."""
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


# FLOP credited per interaction to the fused forward + adjoint step: the
# forward step's 20 plus the adjoint's own arithmetic per interaction in the
# reference's target-by-target form (h = g_j - g_k: 3; r . h: 5; 1/d: 2;
# t3 = -3 (r . h) / d: 2; u = t3 r + h: 6; lambda += K u: 6) = 24.
FUSED_FLOP_PER_INTERACTION = 20 + 24
# FLOP the hardware executes per interaction in the fused step (FMA = 2,
# v_rsq_f32 = 1): target-sharded 42 (the forward's 18 + 24); symmetric, per
# pair term (two interactions): r 3, d 6, d^3 2, rsq 1, 1/d 2, h 3, r . h 5,
# t3 2, u 6, target and reaction sums of K u 12, of r K 12 = 54, i.e. 27.
FUSED_EXECUTED_FLOP_PER_INTERACTION = {"targets": 42, "symmetric": 27,
                                       "symmetric_tiles": 27}


def fwd_adjoint_section(nb, torch, pos, vel, n: int, fwd_ms: float) -> dict:
    import numpy as np

    dev = torch.device("cuda:0")
    params = nb.default_params(num_particles=n)
    P = torch.from_numpy(np.stack(pos)).to(dev)
    V = torch.from_numpy(np.stack(vel)).to(dev)
    N = torch.empty_like(P)
    rng = np.random.default_rng(0)
    lxn = torch.from_numpy(rng.normal(size=(3, n)).astype(np.float32)).to(dev)
    lvn = torch.from_numpy(rng.normal(size=(3, n)).astype(np.float32)).to(dev)
    lx, lv = torch.empty_like(lxn), torch.empty_like(lxn)
    reps = 5
    adj_ms = _events(torch, lambda: nb.adjoint_step(params, P, lxn, lvn, lx, lv), reps)
    fused_ms = _events(torch, lambda: nb.fwd_adjoint_step(params, P, V, N, lxn, lvn, lx, lv),
                       reps)
    inter = float(n) * n
    scheme = nb.adjoint_config_info(nb.adjoint_default_config(n))["scheme"]
    credited = FUSED_FLOP_PER_INTERACTION * inter / (fused_ms * 1e-3) / 1e12
    executed = FUSED_EXECUTED_FLOP_PER_INTERACTION[scheme] * inter / (fused_ms * 1e-3) / 1e12
    return {
        "value": round(inter / (fused_ms * 1e-3) / 1e9, 2),
        "unit": "Ginteractions/s",
        "ms_per_step": round(fused_ms, 4),
        "roofline": {
            "bound": "valu",
            "achieved": round(credited, 2),
            "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(credited / PEAK_FP32_TFLOPS, 4),
            "flop_per_interaction": FUSED_FLOP_PER_INTERACTION,
            "scheme": scheme,
            "executed_flop_per_interaction": FUSED_EXECUTED_FLOP_PER_INTERACTION[scheme],
            "executed_frac": round(executed / PEAK_FP32_TFLOPS, 4),
        },
        "ratio_to_fwd": round(fused_ms / fwd_ms, 3),
        "target_ratio": 3.5,
        "adjoint_only_ms": round(adj_ms, 4),
        "fwd_plus_adjoint_ratio": round((fwd_ms + adj_ms) / fwd_ms, 3),
        "note": "fused = one launch computing the forward step and the adjoint at the same "
                "positions (pair terms shared); ratios are against the plain forward step",
    }


def trajectory_section(nb, torch, pos, vel, n: int, fwd_ms: float, steps=20, every=10) -> dict:
    import numpy as np

    dev = torch.device("cuda:0")
    params = nb.default_params(num_particles=n)
    x0 = torch.from_numpy(np.stack(pos)).to(dev).requires_grad_(True)
    v0 = torch.from_numpy(np.stack(vel)).to(dev).requires_grad_(True)
    # one untimed pass first: the workspace and checkpoint allocations
    xT, vT = nb.trajectory(x0, v0, params, steps, every)
    (xT * xT).sum().backward()
    x0.grad = None
    v0.grad = None
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    xT, vT = nb.trajectory(x0, v0, params, steps, every)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    loss = (xT * xT).sum()
    loss.backward()
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    fwd_pass = (t1 - t0) * 1e3 / steps
    bwd_pass = (t2 - t1) * 1e3 / steps
    return {
        "steps": steps,
        "checkpoint_every": every,
        "forward_pass_ms_per_step": round(fwd_pass, 3),
        "backward_pass_ms_per_step": round(bwd_pass, 3),
        "total_ratio_to_fwd": round((fwd_pass + bwd_pass) / fwd_ms, 3),
        "note": "timed after one untimed pass; backward = segment replay from checkpoints "
                "(forward kernel) + adjoint kernel",
    }


def batched_section(nb, steps: int) -> dict:
    import numpy as np

    sizes = nb.plummer_batch_sizes(64, seed=0)
    params, pos, vel = [], [[], [], []], [[], [], []]
    for k, m in enumerate(sizes):
        p, v = nb.plummer_ic(m, seed=1000 + k, a=1.0, g_mass=1.0)
        params.append(nb.default_params(num_particles=m, G=1.0 / m))
        for c in range(3):
            pos[c].append(p[c])
            vel[c].append(v[c])
    b = nb.Batch(params, tuple(np.concatenate(a) for a in pos),
                 tuple(np.concatenate(a) for a in vel))
    best = b.autotune(2)
    b.step(2)
    graph_ms = b.step(steps) / steps
    b.step_separate(2)
    sep_ms = b.step_separate(steps) / steps
    b.close()
    inter = float(sum(m * m for m in sizes))
    return {
        "systems": len(sizes),
        "particles": int(sum(sizes)),
        "interactions_per_step": inter,
        "steps": steps,
        "value": round(inter / (graph_ms * 1e-3) / 1e9, 2),
        "unit": "Ginteractions/s",
        "ms_per_step": round(graph_ms, 4),
        "separate_launches_ms_per_step": round(sep_ms, 4),
        "separate_launches_value": round(inter / (sep_ms * 1e-3) / 1e9, 2),
        "speedup_vs_separate": round(sep_ms / graph_ms, 3),
        "autotuned_launch": nb.batch_config_info(best),
        "ics": "Plummer a=1, G*M=1 per system (G = 1/N_s), seeds 1000+s; sizes seed 0",
    }


def traffic_from_profile(scheme: str, n: int):
    """HBM bytes per launch of the pair kernel of `scheme` from the committed
    rocprofv3 PMC summary (FETCH_SIZE + WRITE_SIZE, calibrated against a
    copy of known size), if one exists for n."""
    f = ROOT / "profiles" / f"forward_{scheme}_n{n}_hbm_traffic.txt"
    if not f.exists():
        return None
    m = re.search(r"^bytes_per_launch:\s*([0-9.eE+]+)", f.read_text(), re.M)
    return float(m.group(1)) if m else None


# FLOP the hardware executes per interaction (one fused multiply-add = 2):
#   targets:   every pair evaluated for both particles: 3 sub + 3 fma + 2 mul
#              + rsqrt + 3 fma = 18;
#   symmetric (super-tile or tile pairs): each pair evaluated once for two
#              interactions: (3 sub + 3 fma + 2 mul + rsqrt + 6 fma) / 2 = 12.
EXECUTED_FLOP_PER_INTERACTION = {"targets": 18, "symmetric": 12, "symmetric_tiles": 12}


def host_cpu() -> dict:
    """The host's CPU model, its logical CPUs and the threads this process may
    use (the oracle's OpenMP threads follow OMP_NUM_THREADS when set)."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            model = next((ln.split(":", 1)[1].strip() for ln in f
                          if ln.startswith("model name")), None)
    except OSError:
        pass
    return {"cpu_model": model, "nproc": os.cpu_count(),
            "affinity": len(os.sched_getaffinity(0)),
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}


def cpu_config1(nb, orc, frames: int = 10, warm: int = 2) -> dict:
    """The CPU configuration (N = 25600 reference disk, SimParam defaults, 10
    steps) timed the reference driver's way (src/nbody.cpp:83-123): one step
    per frame, wall time per frame, the first `warm` frames discarded, mean
    and population standard deviation of the rest."""
    n = 25600
    pos, vel = nb.disk_ic(n)
    p = orc.params(num_particles=n)
    times = []
    for _ in range(frames):
        t0 = time.perf_counter()
        pos, vel = orc.step_f32(p, pos, vel)
        times.append((time.perf_counter() - t0) * 1e3)
    kept = times[warm:]
    mean = sum(kept) / len(kept)
    std = (sum((t - mean) ** 2 for t in kept) / len(kept)) ** 0.5
    return {"n": n, "steps": frames, "steps_per_frame": 1, "warmup_frames": warm,
            "mean_ms": round(mean, 3), "stddev_ms": round(std, 3),
            "value": round(float(n) * n / (mean * 1e-3) / 1e9, 3), "unit": "Ginteractions/s",
            "line": f"At step {frames} kernel time is {kept[-1]:g} and mean is {mean:g} and "
                    f"stddev is: {std:g}"}


def cpu_baseline(nb, pos, vel, n: int, seconds: float) -> dict:
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
        "host": host_cpu(),
        "config1": cpu_config1(nb, orc),
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

    if args.calibrate_copy:
        return calibrate_copy(args.calibrate_copy)
    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        return 1
    n = args.n
    pos, vel = nb.disk_ic(n)
    params = nb.default_params(num_particles=n, sim_iterations_per_frame=args.warmup)
    sim = nb.Simulator(params, pos, vel)
    # direct launches: capturing a frame of `steps` launches as a hipGraph
    # would happen inside the timed frame
    sim.set_graph(False)
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
            "bound": "valu",
            "achieved": round(achieved, 2),
            "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(achieved / PEAK_FP32_TFLOPS, 4),
            "traffic": traffic_from_profile(cfg["scheme"], n),
            "bound_note": "fp32 vector ALU roof: 157.3 TFLOP/s, MI355X's fp32 vector peak "
                          "(two wave64 FMA issues per SIMD every 2 clocks); the pair loop runs "
                          "on the vector ALUs, not the matrix cores (it is not a contraction)",
            "flop_per_interaction": FLOP_PER_INTERACTION,
            "kernel_ms_per_launch": round(per_launch_ms, 4),
            "executed_frac": round(EXECUTED_FLOP_PER_INTERACTION[cfg["scheme"]] * inter
                                   / (per_launch_ms * 1e-3) / 1e12 / PEAK_FP32_TFLOPS, 4),
        },
    }
    if not args.no_extras:
        line["fwd_adjoint"] = fwd_adjoint_section(nb, torch, pos, vel, n, per_launch_ms)
        line["trajectory"] = trajectory_section(nb, torch, pos, vel, n, per_launch_ms)
        line["batched"] = batched_section(nb, args.batch_steps)
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(nb, pos, vel, n, args.cpu_seconds)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
