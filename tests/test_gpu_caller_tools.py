"""GPU: caller-side tools around the engine.

* ``run_until_converged`` stops on the first check whose relative DFT change
  is below ``tol``; its history is that of the same run checked by hand.
* ``python -m hipfdtd.benchmark`` (median of timed windows, key=value file
  as ``benchmarks/benchmark_vs_meep.py:313-331`` writes).
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import fdtd_oracle as fo

pytestmark = pytest.mark.gpu


def _farfield_sim(case):
    import hipfdtd

    sim = hipfdtd.HIPFDTD(tuple(case["shape"]), case["dl"], npml=case["npml"])
    i0, i1, j0, j1, kk = case["patch"]
    sim.add_source(lambda f: f.add_source_Ex(kk, fo.gaussian_sine(f.t, case["freq"], case["fwidth"]),
                                             i0, i1, j0, j1))  # fmt: skip
    mon = hipfdtd.HIPNear2FarMonitor(sim, case["monitor_center"], case["monitor_size"], case["freq"])
    return sim, mon


def test_run_until_converged_matches_manual_checks(cases):
    case = cases["vacuum_farfield"]
    sim, mon = _farfield_sim(case)
    steps, hist = sim.run_until_converged(mon, tol=1e-3, check_every=50, max_steps=8000)
    print(f"[converge] {steps} steps, last {hist[-1]}", flush=True)
    assert hist[-1][1] < 1e-3 and all(r >= 1e-3 for _, r in hist[:-1])
    assert steps == hist[-1][0] == 50 * len(hist) < 8000
    sim2, mon2 = _farfield_sim(case)
    manual = []
    mon2.snapshot_dft()
    for _ in range(len(hist)):
        sim2.run(50)
        manual.append((sim2.step_num, mon2.dft_relative_change()))
        mon2.snapshot_dft()
    assert [h[0] for h in manual] == [h[0] for h in hist]
    np.testing.assert_allclose([h[1] for h in manual], [h[1] for h in hist], rtol=1e-12)
    with pytest.raises(ValueError):
        sim.run_until_converged(mon, check_every=0)


def test_benchmark_cli_writes_results(tmp_path):
    from hipfdtd import benchmark

    out = tmp_path / "bench.txt"
    r = benchmark.main(["--shape", "64,48,80", "--steps", "20", "--warmup", "2", "--repeats", "3",
                        "--npml", "6", "--out", str(out)])  # fmt: skip
    kv = dict(line.split("=", 1) for line in out.read_text().splitlines())
    assert kv["shape"] == "64x48x80" and kv["repeats"] == "3" and kv["path"] in ("split", "fused")
    assert float(kv["hip_mcups_median"]) > 0 and len(r["mcups"]) == 3
    assert float(kv["hip_mcups_min"]) <= float(kv["hip_mcups_median"]) <= float(kv["hip_mcups_max"])
