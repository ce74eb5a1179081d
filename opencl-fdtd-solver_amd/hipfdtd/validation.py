"""Discrepancy metrics and reports for the HIP engine's validation cases.

The reference grades its solver against Meep with objective metrics on
committed fixtures (``tests/meep_validation/plane_metrics.py:54-141``,
``farfield_metrics.py:23-48``, ``harness.py:84-99``) and writes them as JSON
plus Markdown (``update_plane_baselines.py``, ``update_farfield_baselines.py``).
This module restates those metrics and report writers, and runs the HIP
engine on the same four cases (two mid-plane Ex cases, two far-field cases),
so that the engine can be reported against the reference's OpenCL arrays and
its Meep arrays without either OpenCL or Meep:

* plane checkpoints: Pearson correlation after a least-squares amplitude
  scale, the scale itself, raw and aligned residual energy ratios, peaks and
  their ratio on the non-PML mask, and the best integer lag of the mid-x
  lineouts;
* far field: main-lobe max |Δ dB| after aligning peaks (only angles within
  ``mask_db`` of both peaks) and the endfire null ratio |E(+x)|/|E(+z)|.

``python -m hipfdtd.validation --golden tests/golden --out <dir>`` writes
``hip_discrepancy_report{,_farfield}.json`` and the Markdown twins.
"""

from __future__ import annotations

import argparse
import json
import math
from pathlib import Path
from typing import Any

import numpy as np

PLANE_CASES = ("vacuum_sheet", "dielectric_block")
FARFIELD_CASES = ("vacuum_farfield", "dielectric_sphere_farfield")


def round_metric(x: float) -> float:
    """8 significant digits, the reports' serialisation."""
    return float(f"{float(x):.8g}")


def pml_mask(shape2, npml: int) -> np.ndarray:
    """True off the PML frame of a 2-D (x, z) plane."""
    m = np.ones(tuple(shape2), dtype=bool)
    p = max(0, int(npml))
    if p:
        m[:p, :] = m[-p:, :] = False
        m[:, :p] = m[:, -p:] = False
    return m


def _lineout_lag(a_line: np.ndarray, b_line: np.ndarray) -> int:
    """Lag (of b relative to a) maximising the normalised cross-correlation."""
    if a_line.size < 2 or a_line.std() < 1e-30 or b_line.std() < 1e-30:
        return 0
    za = (a_line - a_line.mean()) / (a_line.std() + 1e-30)
    zb = (b_line - b_line.mean()) / (b_line.std() + 1e-30)
    xc = np.correlate(za, zb, mode="full")
    return int(np.argmax(xc)) - (za.size - 1)


def plane_metrics(a, b, *, npml: int, step: int, labels=("opencl", "meep")) -> dict:
    """Metrics of solver plane ``a`` against comparison plane ``b``."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    if a.shape != b.shape:
        raise ValueError(f"shape mismatch at step {step}: {a.shape} vs {b.shape}")
    m = pml_mask(a.shape, npml)
    av, bv = a[m].ravel(), b[m].ravel()
    ea = float(av @ av) + 1e-30
    eb = float(bv @ bv)
    scale = float(av @ bv) / eb if eb > 0.0 else 0.0
    bs = bv * scale
    raw = float((av - bv) @ (av - bv)) / ea
    aligned = float((av - bs) @ (av - bs)) / ea
    if av.std() < 1e-15 or bs.std() < 1e-15:
        corr = 0.0
    else:
        da, db = av - av.mean(), bs - bs.mean()
        corr = float(da @ db) / math.sqrt(float(da @ da) * float(db @ db))
        corr = corr if math.isfinite(corr) else 0.0
    pa = float(np.max(np.abs(av))) if av.size else 0.0
    pb = float(np.max(np.abs(bv))) if bv.size else 0.0
    ix = a.shape[0] // 2
    sl = slice(npml, a.shape[1] - npml) if npml > 0 else slice(None)
    return {
        "step": int(step),
        "pearson_corr": round_metric(corr),
        "lms_scale": round_metric(scale),
        "raw_residual_energy_ratio": round_metric(raw),
        "aligned_residual_energy_ratio": round_metric(aligned),
        f"peak_{labels[0]}": round_metric(pa),
        f"peak_{labels[1]}": round_metric(pb),
        "peak_ratio": round_metric(pa / pb if pb > 0.0 else float("inf")),
        "mid_x_lag_cells": _lineout_lag(a[ix, sl], b[ix, sl]),
    }


def plane_summary(rows) -> dict:
    """Means over the checkpoints (and the largest |lag|)."""
    if not rows:
        keys = ("mean_pearson_corr", "mean_lms_scale", "mean_raw_residual_energy_ratio",
                "mean_aligned_residual_energy_ratio", "max_abs_mid_x_lag_cells")  # fmt: skip
        return {k: 0.0 for k in keys}
    mean = lambda k: round_metric(float(np.mean([r[k] for r in rows])))  # noqa: E731
    return {
        "mean_pearson_corr": mean("pearson_corr"),
        "mean_lms_scale": mean("lms_scale"),
        "mean_raw_residual_energy_ratio": mean("raw_residual_energy_ratio"),
        "mean_aligned_residual_energy_ratio": mean("aligned_residual_energy_ratio"),
        "max_abs_mid_x_lag_cells": float(max(abs(r["mid_x_lag_cells"]) for r in rows)),
    }


def max_abs_db_error(a_db, b_db, *, mask_db=None) -> float:
    """Max |Δ| after shifting both patterns' peaks to 0 dB, over the angles
    where both lie within ``mask_db`` of their peak (all angles if None or if
    no angle qualifies)."""
    a = np.asarray(a_db, dtype=np.float64)
    b = np.asarray(b_db, dtype=np.float64)
    a0, b0 = a - a.max(), b - b.max()
    if mask_db is not None:
        m = (a0 >= mask_db) & (b0 >= mask_db)
        if m.any():
            return float(np.max(np.abs(a0[m] - b0[m])))
    return float(np.max(np.abs(a0 - b0)))


def null_ratio(eh_main, eh_null) -> float:
    """|E(null)| / |E(main)| of packed (Ex, Ey, Ez, Hx, Hy, Hz) vectors."""
    a = np.asarray(eh_main, dtype=np.complex128)
    b = np.asarray(eh_null, dtype=np.complex128)
    return float(np.linalg.norm(b[:3]) / (np.linalg.norm(a[:3]) + 1e-30))


def farfield_metrics(a_db, b_db, a_eh: dict, b_eh: dict, *, mask_db=-12.0,
                     labels=("opencl", "meep")) -> dict:  # fmt: skip
    return {
        "main_lobe_max_abs_db_error": round_metric(max_abs_db_error(a_db, b_db, mask_db=mask_db)),
        f"{labels[0]}_null_ratio": round_metric(null_ratio(a_eh["eh_plus_z"], a_eh["eh_plus_x"])),
        f"{labels[1]}_null_ratio": round_metric(null_ratio(b_eh["eh_plus_z"], b_eh["eh_plus_x"])),
    }


# ── the HIP engine on the validation cases ──


def _gaussian_sine(t, f0, fw):
    """Gaussian-modulated sine of the reference harness
    (tests/meep_validation/harness.py:70-73)."""
    t0 = 5.0 / (math.pi * fw)
    sig = 1.0 / (math.pi * fw)
    return math.exp(-0.5 * ((t - t0) / sig) ** 2) * math.sin(2.0 * math.pi * f0 * t)


def plane_epsilon(case) -> np.ndarray:
    """Centred dielectric block of a plane case (plane_cases.py:97-109)."""
    eps = np.ones(case["shape"], dtype=np.float32)
    hx, hy, hz = case["block_half"]
    if hx > 0 or hy > 0 or hz > 0:
        cx, cy, cz = (n // 2 for n in case["shape"])
        eps[cx - hx : cx + hx + 1, cy - hy : cy + hy + 1, cz - hz : cz + hz + 1] = case["block_eps"]
    return eps


def sphere_epsilon(case) -> np.ndarray:
    """Staircased sphere of a far-field case (cases.py:86-96)."""
    eps = np.ones(case["shape"], dtype=np.float32)
    if case["sphere_eps"]:
        c = case["shape"][0] // 2
        I, J, K = np.indices(case["shape"])
        eps[(I - c) ** 2 + (J - c) ** 2 + (K - c) ** 2 < int(case["sphere_rad_cells"]) ** 2] = (
            case["sphere_eps"])  # fmt: skip
    return eps


def run_plane_case(case, **engine_kw) -> dict:
    """Mid-plane Ex[:, Ny//2, :] at the checkpoints (plane_cases.py:116-146):
    Jx sheet at z = Nz - npml - 2 over [npml, N - npml)^2, rim taper 0.8."""
    from .engine import HIPFDTD

    nx, ny, nz = case["shape"]
    p = case["npml"]
    sim = HIPFDTD(tuple(case["shape"]), case["dl"], npml=p, **engine_kw)
    sim.set_epsilon(plane_epsilon(case))
    z = nz - p - 2
    sim.add_source(lambda f: f.add_source_Jx(z, _gaussian_sine(f.t, case["freq"], case["fwidth"]),
                                             i0=p, i1=nx - p, j0=p, j1=ny - p, rim_taper=True))  # fmt: skip
    out = {}
    for s in range(1, case["n_steps"] + 1):
        sim.step()
        if s in case["checkpoints"]:
            out[s] = sim.Ex[:, ny // 2, :].copy()
    sim.close()
    return out


def run_farfield_case(case, **engine_kw):
    """XZ far-field pattern (dB) and E/H at R along +z and +x of a compact
    soft-Ex patch (cases.py:23-83, 249-264)."""
    from .engine import HIPFDTD
    from .monitors import HIPNear2FarMonitor

    sim = HIPFDTD(tuple(case["shape"]), case["dl"], npml=case["npml"], **engine_kw)
    if case["sphere_eps"]:
        sim.set_epsilon(sphere_epsilon(case))
    i0, i1, j0, j1, kk = case["patch"]
    sim.add_source(lambda f: f.add_source_Ex(kk, _gaussian_sine(f.t, case["freq"], case["fwidth"]),
                                             i0, i1, j0, j1))  # fmt: skip
    mon = HIPNear2FarMonitor(sim, case["monitor_center"], case["monitor_size"], case["freq"])
    sim.run(case["n_steps"])
    R = case["obs_r"]
    _, db = mon.farfield_polar_xz(distance_m=R, n_angles=case["n_angles"])
    eh = {"eh_plus_z": mon.get_farfield((0.0, 0.0, R)), "eh_plus_x": mon.get_farfield((R, 0.0, 0.0))}
    sim.close()
    return db, eh


def _courant() -> float:
    return 0.99 / math.sqrt(3.0)


def build_reports(golden: Path, against=("opencl", "meep"), **engine_kw):
    """Run the HIP engine on the four cases and compare it with the committed
    arrays of each solver in ``against`` ('opencl' -> ocl_*, 'meep' ->
    meep_*).  Returns (plane document, far-field document)."""
    golden = Path(golden)
    cases = json.loads((golden / "cases.json").read_text())
    prefix = {"opencl": "ocl", "meep": "meep"}
    plane_cases = []
    for name in PLANE_CASES:
        c = cases[name]
        got = run_plane_case(c, **engine_kw)
        entry = {"name": name, "shape": c["shape"], "dl_m": c["dl"], "npml": c["npml"],
                 "n_steps": c["n_steps"], "freq_hz": c["freq"], "fwidth_hz": c["fwidth"],
                 "courant": _courant(), "block_half": c["block_half"],
                 "block_eps": c["block_eps"], "against": {}}  # fmt: skip
        for ref in against:
            rows = [plane_metrics(got[s], np.load(golden / name / f"{prefix[ref]}_ex_step{s:04d}.npy"),
                                  npml=c["npml"], step=s, labels=("hip", ref))
                    for s in c["checkpoints"]]  # fmt: skip
            entry["against"][ref] = {"checkpoints": rows, "summary": plane_summary(rows)}
        plane_cases.append(entry)
    ff_cases = []
    for name in FARFIELD_CASES:
        c = cases[name]
        db, eh = run_farfield_case(c, **engine_kw)
        entry = {"name": name, "shape": c["shape"], "dl_m": c["dl"], "npml": c["npml"],
                 "n_steps": c["n_steps"], "n_angles": c["n_angles"], "freq_hz": c["freq"],
                 "fwidth_hz": c["fwidth"], "courant": _courant(), "sphere_eps": c["sphere_eps"],
                 "sphere_rad_cells": c["sphere_rad_cells"], "against": {}}  # fmt: skip
        for ref in against:
            p = prefix[ref]
            ref_db = np.load(golden / name / f"{p}_S_db.npy")
            ref_eh = {k: np.load(golden / name / f"{p}_{k}.npy") for k in ("eh_plus_z", "eh_plus_x")}
            entry["against"][ref] = farfield_metrics(db, ref_db, eh, ref_eh, labels=("hip", ref))
        ff_cases.append(entry)
    plane_doc = {
        "title": "HIP engine mid-plane Ex discrepancy report",
        "description": ("Mid-plane Ex of the HIP engine against the reference's committed "
                        "OpenCL and Meep arrays on the same cases (SI Jx sheet with rim taper). "
                        "Metrics use the non-PML mask."),  # fmt: skip
        "cases": plane_cases,
    }
    ff_doc = {
        "title": "HIP engine near-to-far discrepancy report",
        "description": ("XZ far-field |S|(θ) and endfire null depth of the HIP engine against "
                        "the reference's committed OpenCL and Meep arrays."),  # fmt: skip
        "cases": ff_cases,
    }
    return plane_doc, ff_doc


def plane_markdown(doc) -> str:
    lines = [f"# {doc['title']}", "", doc["description"], ""]
    for case in doc["cases"]:
        lines += [f"## Case `{case['name']}`", "",
                  f"- Grid `{tuple(case['shape'])}`, dl={case['dl_m']} m, npml={case['npml']}, "
                  f"n_steps={case['n_steps']}; block half-cells={tuple(case['block_half'])}, "
                  f"ε={case['block_eps']}", ""]  # fmt: skip
        for ref, res in case["against"].items():
            s = res["summary"]
            lines += [f"### HIP vs {ref}", "",
                      "| step | corr | LMS scale | raw res/E | aligned res/E | peak ratio | lag |",
                      "|-----:|-----:|----------:|----------:|--------------:|-----------:|----:|"]  # fmt: skip
            for r in res["checkpoints"]:
                lines.append(f"| {r['step']} | {r['pearson_corr']:.6f} | {r['lms_scale']:.6f} | "
                             f"{r['raw_residual_energy_ratio']:.6g} | "
                             f"{r['aligned_residual_energy_ratio']:.6g} | "
                             f"{r['peak_ratio']:.6f} | {r['mid_x_lag_cells']} |")  # fmt: skip
            lines += ["", f"Mean corr {s['mean_pearson_corr']:.6f}, mean LMS scale "
                      f"{s['mean_lms_scale']:.6f}, mean aligned residual "
                      f"{s['mean_aligned_residual_energy_ratio']:.6g}, max |lag| "
                      f"{int(s['max_abs_mid_x_lag_cells'])}.", ""]  # fmt: skip
    return "\n".join(lines)


def farfield_markdown(doc) -> str:
    lines = [f"# {doc['title']}", "", doc["description"], "",
             "| case | vs | main lobe max |Δ| dB | HIP null | other null |",
             "|------|----|------------------:|---------:|-----------:|"]  # fmt: skip
    for case in doc["cases"]:
        for ref, m in case["against"].items():
            lines.append(f"| `{case['name']}` | {ref} | {m['main_lobe_max_abs_db_error']:.6f} | "
                         f"{m['hip_null_ratio']:.6e} | {m[f'{ref}_null_ratio']:.6e} |")  # fmt: skip
    lines.append("")
    return "\n".join(lines)


def write_reports(root, plane_doc, ff_doc):
    """JSON (sorted keys, round-tripped) and Markdown for both documents."""
    root = Path(root)
    root.mkdir(parents=True, exist_ok=True)
    out = []
    for stem, doc, md in (("hip_discrepancy_report", plane_doc, plane_markdown),
                          ("hip_discrepancy_report_farfield", ff_doc, farfield_markdown)):  # fmt: skip
        doc = json.loads(json.dumps(doc, sort_keys=True))
        j = root / f"{stem}.json"
        j.write_text(json.dumps(doc, indent=2, sort_keys=True) + "\n", encoding="utf-8")
        m = root / f"{stem}.md"
        m.write_text(md(doc), encoding="utf-8")
        out += [j, m]
    return out


def main(argv=None) -> None:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--golden", required=True, help="directory holding cases.json and the arrays")
    ap.add_argument("--out", required=True, help="directory for the reports")
    ap.add_argument("--against", default="opencl,meep")
    args = ap.parse_args(argv)
    plane_doc, ff_doc = build_reports(Path(args.golden), tuple(args.against.split(",")))
    for p in write_reports(args.out, plane_doc, ff_doc):
        print(p)


if __name__ == "__main__":
    main()
