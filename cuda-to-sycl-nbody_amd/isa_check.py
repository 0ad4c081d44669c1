"""Build-time scan of the gfx950 assembly of libnbody_amd (host-side tool,
run by build.py; never loaded at run time).

Fails the build when the device code contains
  * packed fp32 VALU instructions (v_pk_add_f32 / v_pk_mul_f32 /
    v_pk_fma_f32): this project's device code is built with them disabled
    and must stay so;
  * any write through the scalar data cache (scalar stores, scalar atomics,
    scalar cache write-back or discard): device code writes memory only
    with vector instructions.
"""
from __future__ import annotations

import re
from pathlib import Path

FORBIDDEN = {
    "packed fp32": re.compile(r"^\s*v_pk_(add|mul|fma)_f32\b", re.M),
    "scalar store": re.compile(r"^\s*s_(buffer_|scratch_)?store", re.M),
    "scalar atomic": re.compile(r"^\s*s_(buffer_)?atomic", re.M),
    "scalar cache write-back/discard": re.compile(r"^\s*s_dcache_(wb|discard)", re.M),
}


def check_dir(asm_dir: Path, arch: str) -> None:
    bad = []
    files = sorted(Path(asm_dir).glob(f"*-hip-amdgcn-amd-amdhsa-{arch}.s"))
    if not files:
        raise RuntimeError(f"no {arch} assembly found in {asm_dir}")
    for s in files:
        text = s.read_text()
        for what, rx in FORBIDDEN.items():
            n = len(rx.findall(text))
            if n:
                bad.append(f"{s.name}: {n} x {what}")
    if bad:
        raise RuntimeError("forbidden instructions in device code: " + "; ".join(bad))
