"""Committed fixtures (data only) and their loaders.

disk_seed5489_n256.txt  first 256 particles of the reference disk generator
                        (src/simulator.cu:131-158, default engine): one line
                        per particle, x y z vx vy vz as float32 bit patterns
                        in hex.
step1_n256.txt          the oracle's fp32 step of those 256 particles with
                        SimParam defaults: x' y' z' vx' vy' vz' per line.
Written by tests/golden/make_fixtures.py.
"""
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent


def _load(name: str) -> np.ndarray:
    rows = [[int(w, 16) for w in line.split()] for line in (HERE / name).read_text().splitlines()
            if line and not line.startswith("#")]
    return np.array(rows, dtype=np.uint32).T.copy().view(np.float32)


def load_disk_fixture() -> np.ndarray:
    """(6, 256) float32: x, y, z, vx, vy, vz."""
    return _load("disk_seed5489_n256.txt")


def load_step1_fixture() -> np.ndarray:
    """(6, 256) float32 after one default-parameter step."""
    return _load("step1_n256.txt")
