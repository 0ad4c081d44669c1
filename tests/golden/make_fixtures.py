"""Writes the fixtures in this folder from the CPU oracle.

    python tests/golden/make_fixtures.py

The oracle's disk generator is pinned bit-for-bit against libstdc++'s
std::mt19937 / uniform_real_distribution (tests/test_oracle.py), so the disk
fixture records the reference generator's output for seed 5489.
"""
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent.parent / "oracle"))

import oracle  # noqa: E402


def _write(name: str, arrs, header: str) -> None:
    words = np.stack([np.asarray(a, np.float32) for a in arrs]).view(np.uint32).T
    lines = [f"# {header}"] + [" ".join(f"{w:08x}" for w in row) for row in words]
    (HERE / name).write_text("\n".join(lines) + "\n")


def main() -> None:
    pos, vel = oracle.disk(256)
    _write("disk_seed5489_n256.txt", pos + vel, "reference disk, seed 5489, n=256: x y z vx vy vz")
    p = oracle.params()
    nxt, nv = oracle.step_f32(p, pos, vel)
    _write("step1_n256.txt", nxt + nv, "one fp32 step, SimParam defaults: x y z vx vy vz")


if __name__ == "__main__":
    main()
