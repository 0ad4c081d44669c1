"""Builds libnbody_amd (HIP kernels + C ABI), libnbody_sim (C++ Simulator
over the C ABI), the nbody_amd driver and the nb_issue_rates
microbenchmark, in-tree under lib/.

    python cuda-to-sycl-nbody_amd/build.py

Every HIP file is compiled for gfx950 only, with -ffp-contract=off (the
integrator rounds as the reference source is written) and with packed fp32
VALU instructions disabled.  After the build the gfx950 assembly is scanned
and the build fails if any packed fp32 instruction was emitted.
"""
from __future__ import annotations

import os
import re
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
INCLUDE = ROOT / "include"
CSRC = PKG / "csrc"
HOST = PKG / "host"
LIB = PKG / "lib"
BUILD = ROOT / "_build"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")
ARCH = "gfx950"

HIP_SOURCES = ["nb_forward.hip", "nb_sym.hip", "nb_adjoint.hip", "nb_traj.hip", "nb_batch.hip", "nb_diag.hip", "nb_pack.hip", "nb_capi.hip"]
HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-fPIC",
    "-std=c++17",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-command-line-argument",
    f"-I{INCLUDE}",
    f"-I{CSRC}",
]
CXX_FLAGS = ["-O2", "-fPIC", "-std=c++17", "-ffp-contract=off", "-Wall", f"-I{INCLUDE}"]



def _run(cmd: list[str], cwd: Path | None = None) -> None:
    print("+", " ".join(str(c) for c in cmd), flush=True)
    res = subprocess.run([str(c) for c in cmd], cwd=cwd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout)
        sys.stderr.write(res.stderr)
        raise RuntimeError(f"build step failed ({res.returncode}): {cmd[0]}")
    # hipcc's host pass warns that it does not know the gfx feature: drop it
    noise = "is not a recognized feature for this target"
    for line in (res.stdout + res.stderr).splitlines():
        if line.strip() and noise not in line:
            print(line)


def build(verbose: bool = True) -> Path:
    """Compiles everything; returns the lib/ directory."""
    if shutil.which(HIPCC) is None and not Path(HIPCC).exists():
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    LIB.mkdir(parents=True, exist_ok=True)
    obj_dir = BUILD / "obj"
    obj_dir.mkdir(parents=True, exist_ok=True)

    objs = []
    for src in HIP_SOURCES:
        obj = obj_dir / (Path(src).stem + ".o")
        # --save-temps writes the gfx950 assembly next to the object (cwd)
        _run([HIPCC, *HIP_FLAGS, "--save-temps", "-c", CSRC / src, "-o", obj], cwd=obj_dir)
        objs.append(obj)
    # the issue-rate microbenchmark of the pair loop's instructions (a program)
    probe = LIB / "nb_issue_rates"
    _run([HIPCC, *HIP_FLAGS, "--save-temps", PKG / "probe" / "issue_rates.hip", "-o", probe],
         cwd=obj_dir)
    import isa_check  # scans the gfx950 assembly written by --save-temps

    isa_check.check_dir(obj_dir, ARCH)

    ic_obj = obj_dir / "ic.o"
    _run([CXX, *CXX_FLAGS, "-c", HOST / "ic.cpp", "-o", ic_obj])
    objs.append(ic_obj)

    lib = LIB / "libnbody_amd.so"
    tmp = LIB / "libnbody_amd.so.tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp])
    os.replace(tmp, lib)

    sim = LIB / "libnbody_sim.so"
    tmp = LIB / "libnbody_sim.so.tmp"
    _run([CXX, *CXX_FLAGS, "-shared", HOST / "sim_param.cpp", HOST / "simulator.cpp",
          f"-L{LIB}", "-lnbody_amd", "-Wl,-rpath,$ORIGIN", "-o", tmp])
    os.replace(tmp, sim)

    drv = LIB / "nbody_amd"
    _run([CXX, *CXX_FLAGS, HOST / "nbody_main.cpp", f"-L{LIB}", "-lnbody_sim", "-lnbody_amd",
          "-Wl,-rpath,$ORIGIN", "-o", drv])
    return LIB


if __name__ == "__main__":
    build()
