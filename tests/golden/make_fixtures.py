"""Copy the reference's committed POCL FP32 output arrays into ``tests/golden/``.

Run once in a container that has the reference checkout at ``/root/reference``
(``python tests/golden/make_fixtures.py``).  Only data files are copied: the
``ocl_*.npy`` outputs and ``angles_deg.npy`` that the reference's own golden
tests compare against (``tests/test_meep_plane_baselines.py:99-121``,
``tests/test_meep_farfield_baselines.py:103-144``), the ``meep_*.npy`` arrays
its discrepancy reports compare them with, and those two committed reports
(``baselines/discrepancy_report{,_farfield}.json``).  They are loaded with
``numpy.load(allow_pickle=False)``.  The inputs that produced them are restated
in ``cases.json`` from the case definitions in
``tests/meep_validation/plane_cases.py:62-146`` and
``tests/meep_validation/cases.py:23-83, 249-264``.
"""

from __future__ import annotations

import json
import shutil
from pathlib import Path

import numpy as np

REF = Path("/root/reference/tests/meep_validation/baselines")
HERE = Path(__file__).resolve().parent

PLANE_CASES = ("vacuum_sheet", "dielectric_block")
FARFIELD_CASES = ("vacuum_farfield", "dielectric_sphere_farfield")

CASES = {
    # plane_cases.py:68-89 — Jx sheet with rim taper, mid-plane Ex[:, Ny//2, :]
    "vacuum_sheet": {
        "kind": "plane",
        "shape": [32, 32, 40],
        "dl": 2.5e-3,
        "npml": 4,
        "n_steps": 80,
        "freq": 5e9,
        "fwidth": 1e9,
        "checkpoints": [20, 40, 60, 80],
        "block_half": [0, 0, 0],
        "block_eps": 1.0,
        "source": "Jx sheet z=Nz-npml-2 over [npml,N-npml)^2, rim_taper, rim_edge 0.8, renorm",
    },
    "dielectric_block": {
        "kind": "plane",
        "shape": [32, 32, 40],
        "dl": 2.5e-3,
        "npml": 4,
        "n_steps": 80,
        "freq": 5e9,
        "fwidth": 1e9,
        "checkpoints": [20, 40, 60, 80],
        "block_half": [6, 6, 4],
        "block_eps": 4.0,
        "source": "Jx sheet z=Nz-npml-2 over [npml,N-npml)^2, rim_taper, rim_edge 0.8, renorm",
    },
    # cases.py:23-83, 249-264 — compact 5x5 soft-Ex patch, N2F box, 19-angle XZ cut
    "vacuum_farfield": {
        "kind": "farfield",
        "shape": [30, 30, 30],
        "dl": 2e-3,
        "npml": 6,
        "n_steps": 500,
        "freq": 5e9,
        "fwidth": 1e9,
        "monitor_center": [30e-3, 30e-3, 30e-3],
        "monitor_size": [20e-3, 20e-3, 20e-3],
        "obs_r": 10.0,
        "n_angles": 19,
        "patch": [13, 18, 13, 18, 15],
        "sphere_eps": None,
        "sphere_rad_cells": 4,
    },
    "dielectric_sphere_farfield": {
        "kind": "farfield",
        "shape": [30, 30, 30],
        "dl": 2e-3,
        "npml": 6,
        "n_steps": 500,
        "freq": 5e9,
        "fwidth": 1e9,
        "monitor_center": [30e-3, 30e-3, 30e-3],
        "monitor_size": [20e-3, 20e-3, 20e-3],
        "obs_r": 10.0,
        "n_angles": 19,
        "patch": [13, 18, 13, 18, 15],
        "sphere_eps": 4.0,
        "sphere_rad_cells": 4,
    },
}


REPORTS = {
    # the reference's committed OpenCL-vs-Meep discrepancy reports (expected
    # outputs of its metric functions on the ocl_* / meep_* arrays)
    "discrepancy_report.json": "reference_discrepancy_report.json",
    "discrepancy_report_farfield.json": "reference_discrepancy_report_farfield.json",
}


def main() -> None:
    for name in PLANE_CASES + FARFIELD_CASES:
        dst = HERE / name
        dst.mkdir(exist_ok=True)
        for src in sorted((REF / name).glob("*.npy")):
            if not (src.name.startswith(("ocl_", "meep_")) or src.name == "angles_deg.npy"):
                continue
            np.load(src, allow_pickle=False)  # data only; refuse anything pickled
            shutil.copyfile(src, dst / src.name)
    for src, dst in REPORTS.items():
        json.loads((REF / src).read_text())  # plain JSON data
        shutil.copyfile(REF / src, HERE / dst)
    (HERE / "cases.json").write_text(json.dumps(CASES, indent=2, sort_keys=True) + "\n")


if __name__ == "__main__":
    main()
