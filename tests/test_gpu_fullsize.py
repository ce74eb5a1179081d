"""Short runs of every configuration at its full size, on the GPU.

* 256^3 benchmark workload: FP32 (the benchmark's path) against the FP64
  engine on the same inputs over 200 steps, for the fields, the 8-frequency
  Huygens-box DFT and the far field at 8 frequencies x 362 angles; every
  result within 1e-4 peak-normalised (the FP64 rerun the configuration names
  as its oracle; reference ``tests/test_unit_engine.py:340-370`` compares
  FP32 and FP64 runs the same way).
* 512^3 plane-DFT workload: the march sweep with two rows per wavefront and
  ce volumes (the engine's own choice on this medium) and three in-sweep
  dipoles, bit-exact against the fused sweep; the K = 32 matrix-core plane
  DFT within 1e-5 of FP64 sums after 64 steps.
* 64 uneven grids (48-160 cells per axis): the batched split sweep bitwise
  against serial split engines over 30 steps; two of the grids against the
  oracle.
* 160^3 adjoint: FP32 against FP64 gradient over 300 steps.
"""

from __future__ import annotations

import numpy as np
import pytest

import bench
from oracle import fdtd_oracle as fo

pytestmark = pytest.mark.gpu

FIELDS = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")


def hip():
    import hipfdtd
    import hipfdtd.batch
    import hipfdtd.dft
    import hipfdtd.probes

    return hipfdtd


def test_256cube_fp32_against_fp64_200_steps():
    print("[c2-fp64] 256^3 benchmark workload, 200 steps, FP32 vs FP64", flush=True)
    res = bench.run_accuracy(bench.Workload(), None, 200)
    print(f"[c2-fp64] {res}", flush=True)
    assert res["fp32_path"] == "march"
    for key in ("fields", "box_dft", "farfield"):
        for grp, err in res[key].items():
            assert err < 1e-4, (key, grp, err)


def _c3_engine(path, eps, pos, f0, fw):
    sim = hip().HIPFDTD((512, 512, 512), 1e-3, npml=16, path=path)
    sim.set_epsilon(eps)
    for comp, p, lag in zip("xyz", pos, (0, 20, 40)):
        box = (p[0], p[0] + 1, p[1], p[1] + 1, p[2], p[2] + 1)
        sim.add_source(lambda f, comp=comp, box=box, lag=lag:
                       f.add_source_J(comp, bench.gaussian_sine(f.t - lag * f.dt, f0, fw), box))  # fmt: skip
    return sim


def test_512cube_march_two_rows_matches_fused_and_plane_dft():
    shape, npml = (512, 512, 512), 16
    f0 = bench.C0 / (20 * 1e-3)
    fw = 0.5 * f0
    freqs = np.linspace(f0 - fw / 2, f0 + fw / 2, 16)
    eps = bench.c3_epsilon(shape)
    rng = np.random.default_rng(3)
    pos = [tuple(int(rng.integers(npml + 16, s - npml - 16)) for s in shape) for _ in range(3)]
    print(f"[c3] 512^3 smooth medium, dipoles at {pos}", flush=True)

    ref = _c3_engine("fused", eps, pos, f0, fw)
    ref.run(4)
    want = ref.fields()
    ref.close()

    sim = _c3_engine("march", eps, pos, f0, fw)
    assert sim.material_count() == 0  # too many distinct values: the ce volumes
    mons = {m: [hip().dft.HIPPlaneDFTMonitor(sim, ax, 256, ("Ex", "Ey", "Ez"), freqs, **kw)
                for ax in "xyz"]
            for m, kw in (("buffered", dict(mode="buffered", K=32)),
                          ("fp64", dict(mode="direct", accum="fp64")))}  # fmt: skip
    sim.run(4)
    got = sim.fields()
    for n in FIELDS:
        np.testing.assert_array_equal(got[n], want[n], err_msg=n)
    del got, want
    sim.run(60)
    print("[c3] march == fused after 4 steps; plane DFT after 64 steps", flush=True)
    checked = 0
    for a, b in zip(mons["buffered"], mons["fp64"]):
        r = b.dft()
        if np.max(np.abs(r)) == 0:  # the pulse has not reached this plane
            continue
        err = fo.peak_normalised_error(a.dft(), r)
        print(f"[c3] plane {a.axis}: buffered vs fp64 {err:.2e}", flush=True)
        assert err < 1e-5, err
        checked += 1
    assert checked >= 1
    sim.close()


def _c4_batch_and_serial(cfg, dl, npml, steps):
    """Batched split sweep and serial split engines on the grids of ``cfg``
    (the benchmark's source and probe for each); returns the batch's sims,
    its probe DFTs and the serial (fields, probe DFT) pairs."""

    def wave(c):
        return lambda t: bench.gaussian_sine(t, c["f"], 0.5 * c["f"])

    def box(c):
        x, y, z = c["centre"]
        return (x, x + 1, y, y + 1, z, z + 1)

    serial = []
    for c in cfg:
        sim = hip().HIPFDTD(c["shape"], dl, npml=npml, path="split")
        sim.set_epsilon(bench.c4_eps(c))
        w, bx = wave(c), box(c)
        sim.add_source(lambda f, w=w, bx=bx: f.add_source_J("z", w(f.t), bx))
        pr = hip().probes.HIPPointProbe(sim, c["probe"], "Ez", c["freqs"])
        sim.run(steps)
        serial.append((sim.fields(), pr.dft()))
        sim.close()
    sims = []
    for c in cfg:
        sim = hip().HIPFDTD(c["shape"], dl, npml=npml)
        sim.set_epsilon(bench.c4_eps(c))
        sims.append(sim)
    batch = hip().batch.BatchedFDTD(sims)
    batch.set_path("split")
    batch.set_sources([("z", box(c)) for c in cfg], [wave(c) for c in cfg], steps)
    batch.set_probes([c["probe"] for c in cfg], "Ez", [c["freqs"] for c in cfg], steps)
    batch.run(steps, graph_steps=10)
    dfts = batch.probe_dfts()
    return batch, sims, dfts, serial


def test_64_uneven_grids_batch_matches_serial_bitwise():
    cfg, dl, npml = bench.c4_sims()
    cells = sum(int(np.prod(c["shape"])) for c in cfg)
    print(f"[c4] 64 grids, {cells} cells, 30 steps", flush=True)
    batch, sims, dfts, serial = _c4_batch_and_serial(cfg, dl, npml, 30)
    for q, sim in enumerate(sims):
        ref_f, ref_p = serial[q]
        got = sim.fields()
        for n in FIELDS:
            np.testing.assert_array_equal(got[n], ref_f[n], err_msg=f"grid {q} {n}")
        np.testing.assert_array_equal(dfts[q], ref_p, err_msg=f"grid {q} probe")
        assert float(np.max(np.abs(got["Ez"]))) > 0, q
    batch.close()
    for s in sims:
        s.close()


def test_batched_grids_match_oracle():
    """The two smallest of the 64 grids, batched, against the oracle: fields
    and the probe DFT (a host DFT of the oracle's Ez at the probe cell)."""
    cfg, dl, npml = bench.c4_sims()
    small = sorted(cfg, key=lambda c: int(np.prod(c["shape"])))[:2]
    steps = 120
    print(f"[c4-oracle] shapes {[c['shape'] for c in small]}, {steps} steps", flush=True)
    batch, sims, dfts, _ = _c4_batch_and_serial(small, dl, npml, steps)
    for q, c in enumerate(small):
        ora = fo.OracleFDTD(c["shape"], dl, npml=npml)
        ora.set_epsilon(bench.c4_eps(c))
        x, y, z = c["centre"]
        ora.add_source(lambda f, c=c, b=(x, x + 1, y, y + 1, z, z + 1): f.add_source_J(
            "z", bench.gaussian_sine(f.t, c["f"], 0.5 * c["f"]), b))  # fmt: skip
        om = 2 * np.pi * np.asarray(c["freqs"])
        acc = np.zeros(len(om), complex)
        for _ in range(steps):
            ora.step()
            acc += float(ora.Ez[c["probe"]]) * np.exp(1j * om * ora.t) * ora.dt
        got, ref = sims[q].fields(), ora.fields()
        for grp in (FIELDS[:3], FIELDS[3:]):
            peak = max(float(np.max(np.abs(ref[n]))) for n in grp)
            for n in grp:
                err = float(np.max(np.abs(got[n].astype(np.float64) - ref[n]))) / peak
                assert err < 1e-5, (q, n, err)
        assert np.max(np.abs(dfts[q] - acc)) / np.max(np.abs(acc)) < 1e-5, q
    batch.close()


def test_160cube_adjoint_fp32_against_fp64():
    from hipfdtd.adjoint import ProbeDFTAdjoint

    shape, npml, dl = (160, 160, 160), 12, 1e-3
    f0 = bench.C0 / (30 * dl)
    fw = 0.5 * f0
    rng = np.random.default_rng(5)
    box = (48, 112, 48, 112, 64, 96)
    design = rng.uniform(1.0, 12.0, (64, 64, 32))
    src = [("x", (80, 81, 80, 81, 58, 59), lambda t: bench.gaussian_sine(t, f0, fw))]
    probes = [(0, i, j, 98) for i in range(78, 82) for j in range(78, 82)]
    print("[c5] 160^3 adjoint, 300 steps, checkpoint 100, FP32 vs FP64", flush=True)
    out = {}
    for dt in (np.float32, np.float64):
        pb = ProbeDFTAdjoint(shape, dl, npml, box, src, probes, [f0], 300, checkpoint=100,
                             dtype=dt)  # fmt: skip
        out[np.dtype(dt).name] = pb.value_and_grad(design)
        pb.close()
    (j32, g32), (j64, g64) = out["float32"], out["float64"]
    err = fo.peak_normalised_error(g32, g64)
    print(f"[c5] J rel {abs(j32 - j64) / j64:.2e}, grad peak-normalised {err:.2e}", flush=True)
    assert abs(j32 - j64) / j64 < 1e-4
    assert err < 1e-4
    assert np.max(np.abs(g64)) > 0
