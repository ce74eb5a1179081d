"""GPU: the HIP engine on the reference's validation cases, reported with
the restated discrepancy metrics.

Against the committed OpenCL arrays the HIP engine must be the same solver:
correlation 1 - 1e-9, aligned residual energy < 1e-8, lag 0, far-field main
lobe within 1e-3 dB.  Against the Meep arrays it must show the same
discrepancy the reference's OpenCL solver shows (its committed report):
correlation and LMS scale within 1e-4, main-lobe error within 1e-2 dB, null
ratio within 1 %.  The report files are written and round-trip as JSON.
"""

from __future__ import annotations

import json

import pytest
from conftest import GOLDEN

pytestmark = pytest.mark.gpu


def test_hip_validation_reports(tmp_path):
    from hipfdtd import validation as V

    print("[validation] building reports", flush=True)
    plane_doc, ff_doc = V.build_reports(GOLDEN)
    ref_plane = {c["name"]: c for c in json.loads(
        (GOLDEN / "reference_discrepancy_report.json").read_text())["cases"]}  # fmt: skip
    ref_ff = {c["name"]: c for c in json.loads(
        (GOLDEN / "reference_discrepancy_report_farfield.json").read_text())["cases"]}  # fmt: skip
    for case in plane_doc["cases"]:
        name = case["name"]
        for r in case["against"]["opencl"]["checkpoints"]:
            assert r["pearson_corr"] > 1 - 1e-9, (name, r)
            assert r["aligned_residual_energy_ratio"] < 1e-8, (name, r)
            assert r["mid_x_lag_cells"] == 0
        ours = case["against"]["meep"]["checkpoints"]
        for a, b in zip(ours, ref_plane[name]["checkpoints"]):
            assert abs(a["pearson_corr"] - b["pearson_corr"]) < 1e-4, (name, a, b)
            assert abs(a["lms_scale"] - b["lms_scale"]) < 1e-4, (name, a, b)
            assert a["mid_x_lag_cells"] == b["mid_x_lag_cells"]
    for case in ff_doc["cases"]:
        name = case["name"]
        assert case["against"]["opencl"]["main_lobe_max_abs_db_error"] < 1e-3, name
        m, r = case["against"]["meep"], ref_ff[name]["metrics"]
        assert abs(m["main_lobe_max_abs_db_error"] - r["main_lobe_max_abs_db_error"]) < 1e-2
        assert abs(m["hip_null_ratio"] / r["opencl_null_ratio"] - 1.0) < 1e-2, (name, m, r)
        print(f"[validation] {name}: {m}", flush=True)
    paths = V.write_reports(tmp_path, plane_doc, ff_doc)
    assert len(paths) == 4
    back = json.loads(paths[0].read_text())
    assert [c["name"] for c in back["cases"]] == list(V.PLANE_CASES)
