"""Shared pytest setup: import paths, the ``gpu`` marker, golden case data."""

from __future__ import annotations

import json
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
PKG_DIR = ROOT / "opencl-fdtd-solver_amd"
GOLDEN = Path(__file__).resolve().parent / "golden"
for p in (str(ROOT), str(PKG_DIR)):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP GPU (MI355X / gfx950)")


@pytest.fixture(scope="session")
def cases():
    return json.loads((GOLDEN / "cases.json").read_text())


def golden(name: str, fname: str) -> np.ndarray:
    return np.load(GOLDEN / name / fname, allow_pickle=False)


def plane_eps(case) -> np.ndarray:
    """ε of a plane case: a centred block (plane_cases.py:97-109)."""
    eps = np.ones(case["shape"], dtype=np.float32)
    hx, hy, hz = case["block_half"]
    if hx > 0 or hy > 0 or hz > 0:
        nx, ny, nz = case["shape"]
        ci, cj, ck = nx // 2, ny // 2, nz // 2
        eps[ci - hx : ci + hx + 1, cj - hy : cj + hy + 1, ck - hz : ck + hz + 1] = case["block_eps"]
    return eps


def sphere_eps(case) -> np.ndarray:
    """Staircased sphere of a far-field case (cases.py:86-96)."""
    eps = np.ones(case["shape"], dtype=np.float32)
    if case["sphere_eps"]:
        ctr = case["shape"][0] // 2
        r2 = int(case["sphere_rad_cells"]) ** 2
        I, J, K = np.indices(case["shape"])
        eps[(I - ctr) ** 2 + (J - ctr) ** 2 + (K - ctr) ** 2 < r2] = case["sphere_eps"]
    return eps
