"""CPU: the restated discrepancy metrics reproduce the reference's own reports.

``hipfdtd.validation`` restates the reference harness's plane and far-field
metrics (tests/meep_validation/plane_metrics.py:54-141,
farfield_metrics.py:23-48, harness.py:84-99).  Applied to the reference's
committed OpenCL and Meep arrays they must give the numbers of its committed
reports (``tests/golden/reference_discrepancy_report{,_farfield}.json``) to
the reports' 8 significant digits (2e-7 relative).  Planted changes (a lineout
shift, a sign flip, a scaled copy) must move the matching metric.
"""

from __future__ import annotations

import json

import numpy as np
import pytest
from conftest import GOLDEN

from hipfdtd import validation as V


def close(a, b, rel=2e-7):
    return abs(a - b) <= rel * max(abs(a), abs(b)) + 1e-30


@pytest.mark.parametrize("name", V.PLANE_CASES)
def test_plane_metrics_match_reference_report(name):
    report = json.loads((GOLDEN / "reference_discrepancy_report.json").read_text())
    case = next(c for c in report["cases"] if c["name"] == name)
    rows = []
    for ref in case["checkpoints"]:
        s = ref["step"]
        got = V.plane_metrics(np.load(GOLDEN / name / f"ocl_ex_step{s:04d}.npy"),
                              np.load(GOLDEN / name / f"meep_ex_step{s:04d}.npy"),
                              npml=case["npml"], step=s)  # fmt: skip
        rows.append(got)
        for k, v in ref.items():
            assert close(got[k], v), (name, s, k, got[k], v)
    summ = V.plane_summary(rows)
    for k, v in case["summary"].items():
        assert close(summ[k], v), (name, k, summ[k], v)


@pytest.mark.parametrize("name", V.FARFIELD_CASES)
def test_farfield_metrics_match_reference_report(name):
    report = json.loads((GOLDEN / "reference_discrepancy_report_farfield.json").read_text())
    case = next(c for c in report["cases"] if c["name"] == name)
    ld = lambda p, k: np.load(GOLDEN / name / f"{p}_{k}.npy")  # noqa: E731
    eh = {p: {k: ld(p, k) for k in ("eh_plus_z", "eh_plus_x")} for p in ("ocl", "meep")}
    got = V.farfield_metrics(ld("ocl", "S_db"), ld("meep", "S_db"), eh["ocl"], eh["meep"])
    for k, v in case["metrics"].items():
        assert close(got[k], v), (name, k, got[k], v)


def test_metrics_respond_to_planted_changes():
    a = np.load(GOLDEN / "vacuum_sheet" / "ocl_ex_step0080.npy").astype(np.float64)
    same = V.plane_metrics(a, a, npml=4, step=80)
    assert same["pearson_corr"] == 1.0 and same["lms_scale"] == 1.0
    assert same["aligned_residual_energy_ratio"] == 0.0 and same["mid_x_lag_cells"] == 0
    half = V.plane_metrics(a, 2.0 * a, npml=4, step=80)
    assert half["lms_scale"] == 0.5 and half["aligned_residual_energy_ratio"] < 1e-12
    flip = V.plane_metrics(a, -a, npml=4, step=80)
    assert flip["lms_scale"] == -1.0 and flip["raw_residual_energy_ratio"] == 4.0
    z = np.arange(40.0)
    bump = np.exp(-0.5 * ((z - 18.0) / 2.0) ** 2)[None, :] * np.ones((32, 1))
    moved = np.exp(-0.5 * ((z - 21.0) / 2.0) ** 2)[None, :] * np.ones((32, 1))
    assert V.plane_metrics(bump, moved, npml=4, step=80)["mid_x_lag_cells"] == -3
    db = np.array([-3.0, 0.0, -1.0, -40.0])
    assert V.max_abs_db_error(db, db + 5.0) == 0.0
    assert V.max_abs_db_error(db, np.array([-3.0, 0.0, -1.0, -10.0])) == 30.0
    assert V.max_abs_db_error(db, np.array([-3.0, 0.0, -1.0, -10.0]), mask_db=-12.0) == 0.0
    eh = np.array([1, 0, 0, 0, 0, 0], complex)
    assert V.null_ratio(eh, 0.1 * eh) == pytest.approx(0.1)
