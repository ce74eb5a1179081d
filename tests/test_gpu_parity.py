"""GPU parity: the HIP engine against the reference fixtures and the oracle.

Tolerances (peak-normalised, max|a-b|/max|b|):
* FP32 HIP vs POCL FP32 fixtures: 1e-4 (target stated for the path);
* FP32 HIP vs the FP32 oracle on the same inputs: 1e-5;
* FP64 HIP vs the FP64 oracle: 1e-12;
* fused sweep vs split sweeps: 1e-6; fused tilings among themselves: bit-exact;
* march sweep vs fused sweep, and march tilings among themselves: bit-exact.
"""

from __future__ import annotations

import numpy as np
import pytest
from conftest import golden, plane_eps, sphere_eps

from oracle import fdtd_oracle as fo
from oracle.n2f_oracle import OracleNear2Far

pytestmark = pytest.mark.gpu

FIELDS = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")


def hip():
    import hipfdtd

    return hipfdtd


def group_error(a: dict, b: dict) -> float:
    """Worst field error normalised by the peak of its E or H group."""
    worst = 0.0
    for grp in (FIELDS[:3], FIELDS[3:]):
        peak = max(float(np.max(np.abs(b[n]))) for n in grp)
        if peak == 0.0:
            continue
        for n in grp:
            worst = max(worst, float(np.max(np.abs(a[n].astype(np.float64) - b[n]))) / peak)
    return worst


def run_plane(sim, case):
    nx, ny, nz = case["shape"]
    p = case["npml"]
    sim.set_epsilon(plane_eps(case))
    z = nz - p - 2
    sim.add_source(
        lambda f: f.add_source_Jx(z, fo.gaussian_sine(f.t, case["freq"], case["fwidth"]),
                                  i0=p, i1=nx - p, j0=p, j1=ny - p, rim_taper=True)
    )  # fmt: skip
    out = {}
    for s in range(1, case["n_steps"] + 1):
        sim.step()
        if s in case["checkpoints"]:
            out[s] = sim.Ex[:, ny // 2, :].copy()
    return out


@pytest.mark.parametrize("name", ["vacuum_sheet", "dielectric_block"])
@pytest.mark.parametrize("path", ["fused", "split", "march"])
def test_planes_match_fixtures_and_oracle(cases, name, path):
    case = cases[name]
    print(f"[plane] {name} path={path}", flush=True)
    sim = hip().HIPFDTD(tuple(case["shape"]), case["dl"], npml=case["npml"], path=path)
    got = run_plane(sim, case)
    ora = run_plane(fo.OracleFDTD(tuple(case["shape"]), case["dl"], npml=case["npml"]), case)
    for s, plane in got.items():
        ref = golden(name, f"ocl_ex_step{s:04d}.npy")
        assert fo.peak_normalised_error(plane, ref) < 1e-4, (s, "fixture")
        assert fo.peak_normalised_error(plane, ora[s]) < 1e-5, (s, "oracle")


def test_fp64_engine_matches_fp64_oracle(cases):
    case = cases["dielectric_block"]
    shp, dl, p = tuple(case["shape"]), case["dl"], case["npml"]
    sim = hip().HIPFDTD(shp, dl, npml=p, dtype=np.float64)
    ora = fo.OracleFDTD(shp, dl, npml=p, dtype=np.float64)
    a = run_plane(sim, case)
    b = run_plane(ora, case)
    for s in a:
        assert fo.peak_normalised_error(a[s], b[s]) < 1e-12, s
    assert group_error(sim.fields(), ora.fields()) < 1e-12


def _random_media(shape, seed, lossy):
    rng = np.random.default_rng(seed)
    eps = rng.uniform(1.0, 6.0, size=shape).astype(np.float32)
    sig = rng.uniform(0.0, 2.0, size=shape).astype(np.float32) if lossy else None
    return eps, sig


def _drive(sim, shape, steps):
    nx, ny, nz = shape
    zc = nz // 2

    def src(f):
        a = fo.gaussian_sine(f.t, 3e10, 1.5e10)
        f.add_source_Jx(zc, a, 2, nx - 3, 3, ny - 2, rim_taper=True, rim_edge=0.7)
        f.add_source_J("z", 0.5 * a, (nx // 3, nx // 3 + 1, ny // 2, ny // 2 + 1, 5, 6))
        f.add_source_J("y", -0.25 * a, (nx - 4, nx - 3, 2, 3, nz - 3, nz - 2))

    sim.add_source(src)
    for _ in range(steps):
        sim.step()


@pytest.mark.parametrize(
    "shape,npml,lossy",
    [((37, 29, 70), 5, False), ((33, 41, 130), 6, True), ((20, 18, 16), 0, False),
     ((24, 24, 64), 4, True)],
)  # fmt: skip
def test_fused_matches_split_and_oracle(shape, npml, lossy):
    print(f"[fused-vs-split] shape={shape} npml={npml} lossy={lossy}", flush=True)
    eps, sig = _random_media(shape, 11, lossy)
    out = {}
    for path in ("fused", "split", "march"):
        sim = hip().HIPFDTD(shape, 1e-3, npml=npml, path=path)
        sim.set_epsilon(eps, sig)
        _drive(sim, shape, 40)
        out[path] = sim.fields()
    ora = fo.OracleFDTD(shape, 1e-3, npml=npml)
    ora.set_epsilon(eps, sig)
    _drive(ora, shape, 40)
    assert group_error(out["fused"], out["split"]) < 1e-6
    for path in ("fused", "split", "march"):
        assert group_error(out[path], ora.fields()) < 1e-5, path
    for n in FIELDS:
        np.testing.assert_array_equal(out["march"][n], out["fused"][n], err_msg=n)


@pytest.mark.parametrize("rows,planes,xcd,lossy", [
    (1, 1, 8, True), (2, 5, 1, True), (4, 16, 8, True), (8, 7, 3, True), (6, 1000, 8, True),
    (10, 9, 8, True), (12, 13, 8, True), (12, 11, 8, False), (10, -1, 2, False)])  # fmt: skip
def test_fused_tilings_are_bit_identical(rows, planes, xcd, lossy):
    shape = (45, 35, 100)
    eps, sig = _random_media(shape, 3, lossy)
    ref = hip().HIPFDTD(shape, 1e-3, npml=6, fused=True)
    ref.set_epsilon(eps, sig)
    _drive(ref, shape, 12)
    sim = hip().HIPFDTD(shape, 1e-3, npml=6, fused=True)
    sim.set_fused_tiling(rows, planes, xcd)
    sim.set_epsilon(eps, sig)
    _drive(sim, shape, 12)
    a, b = ref.fields(), sim.fields()
    for n in FIELDS:
        np.testing.assert_array_equal(a[n], b[n], err_msg=n)


@pytest.mark.parametrize("rows,planes,run,nt,lossy", [
    (2, -1, 16, -1, False), (1, 2, 1, 0, True), (2, 4, 7, 1, True), (1, 1000, 3, 1, False),
    (2, 6, 1, 0, False), (1, -1, 64, 1, True)])  # fmt: skip
def test_march_tilings_are_bit_identical(rows, planes, run, nt, lossy):
    print(f"[march-tiling] rows={rows} planes={planes} run={run} nt={nt} lossy={lossy}",
          flush=True)  # fmt: skip
    shape = (45, 35, 100)
    eps, sig = _random_media(shape, 3, lossy)
    ref = hip().HIPFDTD(shape, 1e-3, npml=6, path="fused")
    ref.set_epsilon(eps, sig)
    _drive(ref, shape, 12)
    sim = hip().HIPFDTD(shape, 1e-3, npml=6, path="march")
    sim.set_march_tiling(rows, planes, run, nt)
    sim.set_epsilon(eps, sig)
    _drive(sim, shape, 12)
    a, b = ref.fields(), sim.fields()
    for n in FIELDS:
        np.testing.assert_array_equal(a[n], b[n], err_msg=n)


def test_set_epsilon_coefficients_match_oracle_bitwise():
    shape = (13, 11, 17)
    eps, sig = _random_media(shape, 5, True)
    sim = hip().HIPFDTD(shape, 1e-3, npml=2)
    sim.set_epsilon(eps)
    for c, ref in zip(("ce_x", "ce_y", "ce_z"), fo.edge_ce(eps, sim.dt)):
        np.testing.assert_array_equal(sim.read_volume(c), ref, err_msg=c)
    sim.set_epsilon(eps, sig)
    ca, cb = fo.edge_ca_cb(eps, sig, sim.dt)
    for c, ref in zip(("ca_x", "ca_y", "ca_z", "ce_x", "ce_y", "ce_z"), ca + cb):
        np.testing.assert_array_equal(sim.read_volume(c), ref, err_msg=c)
    with pytest.raises(ValueError, match="Epsilon shape mismatch"):
        sim.set_epsilon(np.ones((3, 3, 3), np.float32))
    with pytest.raises(ValueError, match="Sigma shape mismatch"):
        sim.set_epsilon(eps, np.ones((3, 3, 3), np.float32))


@pytest.mark.parametrize("path", ["fused", "split", "march"])
def test_jx_closed_form_and_rim(path):
    shape, npml, z = (16, 16, 16), 2, 8
    p = npml
    sim = hip().HIPFDTD(shape, 1e-3, npml=npml, path=path)
    eps = np.ones(shape, np.float32)
    eps[8, 8, z] = 4.0
    sim.set_epsilon(eps)
    sim.add_source(lambda f: f.add_source_Jx(z, 0.25, i0=p, i1=16 - p, j0=p, j1=16 - p))
    sim.step()
    ex = sim.Ex
    assert ex[8, 8, z] == pytest.approx(-sim.dt / (fo.EPS0 * 2.5) * 0.25, rel=1e-6)
    assert ex[p, p, z] == pytest.approx(-sim.dt / fo.EPS0 * 0.25, rel=1e-6)
    assert ex[0, 0, z] == 0.0
    sim2 = hip().HIPFDTD(shape, 1e-3, npml=npml, path=path)
    sim2.add_source(lambda f: f.add_source_Jx(z, 1.0, i0=p, i1=14, j0=p, j1=14, rim_taper=True,
                                              rim_edge=0.5, rim_renorm=False))  # fmt: skip
    sim2.step()
    ex = sim2.Ex
    base = -sim2.dt / fo.EPS0
    assert ex[p + 2, p + 2, z] == pytest.approx(base, rel=1e-6)
    assert ex[p, p + 2, z] == pytest.approx(0.5 * base, rel=1e-6)
    assert ex[p + 2, p, z] == pytest.approx(0.5 * base, rel=1e-6)
    assert ex[p, p, z] == pytest.approx(0.25 * base, rel=1e-6)
    assert ex[13, 13, z] == pytest.approx(0.25 * base, rel=1e-6)


def test_field_access_read_point_and_shapes():
    shape = (14, 15, 16)
    sim = hip().HIPFDTD(shape, 1e-3, npml=3)
    for n in FIELDS:
        arr = getattr(sim, n)
        assert arr.shape == shape and arr.dtype == np.float32
    sim.add_source(lambda f: f.add_source_Ex(5, 0.37))
    sim.step()
    assert sim.read_point("Ex", 6, 7, 5) == float(sim.Ex[6, 7, 5])
    assert sim.step_num == 1 and sim.t == pytest.approx(sim.dt)


@pytest.mark.parametrize("path", ["fused", "split", "march"])
def test_callback_that_edits_fields_matches_soft_source(path):
    """The harness pattern `ex = f.Ex; ex[...] += amp; upload` (cases.py:64-75)
    equals add_source_Ex; the callback sees E before and H after the H update."""
    shape = (24, 22, 26)
    seen = []

    def edit(f):
        ex = f.Ex
        seen.append((float(np.abs(f.Hy).max()), f.t))
        ex[10:15, 9:14, 12] += np.float32(fo.gaussian_sine(f.t, 3e10, 1e10))
        f.Ex = ex

    a = hip().HIPFDTD(shape, 1e-3, npml=4, path=path)
    a.add_source(edit)
    b = hip().HIPFDTD(shape, 1e-3, npml=4, path=path)
    b.add_source(lambda f: f.add_source_Ex(12, fo.gaussian_sine(f.t, 3e10, 1e10), 10, 15, 9, 14))
    for _ in range(30):
        a.step()
        b.step()
    assert group_error(a.fields(), b.fields()) < 1e-6
    assert seen[1][1] == pytest.approx(1.5 * a.dt)


@pytest.mark.parametrize("path", ["fused", "split", "march"])
def test_npml_zero_runs_and_matches_oracle(path):
    shape = (20, 20, 20)
    sim = hip().HIPFDTD(shape, 1e-3, npml=0, path=path)
    ora = fo.OracleFDTD(shape, 1e-3, npml=0)
    for s in (sim, ora):
        s.add_source(lambda f: f.add_source_Ex(10, 0.1))
        for _ in range(5):
            s.step()
    assert sim.psi_x_size == 0
    assert np.all(np.isfinite(sim.Ex)) and float(np.max(np.abs(sim.Ex))) > 0
    assert group_error(sim.fields(), ora.fields()) < 1e-5


def test_sigma_zero_matches_lossless_and_loss_attenuates():
    shape = (24, 24, 60)

    def run(sigma_val):
        sim = hip().HIPFDTD(shape, 1e-3, npml=8)
        eps = np.ones(shape, np.float32)
        sig = np.zeros(shape, np.float32)
        sig[:, :, 40:] = sigma_val
        sim.set_epsilon(eps, sig if sigma_val is not None else None)
        sim.add_source(lambda f: f.add_source_Ex(15, 0.1 * np.sin(2 * np.pi * 10e9 * f.t)))
        sim.run(400)
        return sim.Ex

    np.testing.assert_allclose(run(0.0), run(None), atol=1e-12)
    inside = [abs(run(s)[12, 12, 55]) for s in (20.0, 100.0, 300.0)]
    assert inside[0] > inside[1] > inside[2]


def run_farfield(case, accum=None, freqs=None, fused=True, mon_freq=None, buffer_steps=None):
    sim = hip().HIPFDTD(tuple(case["shape"]), case["dl"], npml=case["npml"], fused=fused)
    if case["sphere_eps"]:
        sim.set_epsilon(sphere_eps(case))
    i0, i1, j0, j1, kk = case["patch"]
    sim.add_source(
        lambda f: f.add_source_Ex(kk, fo.gaussian_sine(f.t, case["freq"], case["fwidth"]),
                                  i0, i1, j0, j1)
    )  # fmt: skip
    mon = hip().HIPNear2FarMonitor(sim, case["monitor_center"], case["monitor_size"],
                                   mon_freq or case["freq"], freqs=freqs, accum=accum,
                                   buffer_steps=buffer_steps)  # fmt: skip
    sim.run(case["n_steps"])
    return sim, mon


@pytest.mark.parametrize("name", ["vacuum_farfield", "dielectric_sphere_farfield"])
def test_farfield_matches_fixtures_and_oracle(cases, name):
    case = cases[name]
    print(f"[farfield] {name}", flush=True)
    sim, mon = run_farfield(case)
    R = case["obs_r"]
    ang, db = mon.farfield_polar_xz(distance_m=R, n_angles=case["n_angles"])
    np.testing.assert_allclose(ang, golden(name, "angles_deg.npy"), atol=1e-12)
    assert np.max(np.abs(db - golden(name, "ocl_S_db.npy"))) < 1e-3
    ez = mon.get_farfield((0.0, 0.0, R))
    ex = mon.get_farfield((R, 0.0, 0.0))
    assert fo.peak_normalised_error(ez, golden(name, "ocl_eh_plus_z.npy")) < 1e-4
    assert fo.peak_normalised_error(ex, golden(name, "ocl_eh_plus_x.npy")) < 1e-4
    # face DFT vs the FP32 oracle on the same inputs
    _, omon = __import__("test_oracle_fixtures").run_farfield_case(case)
    for c in FIELDS:
        got = mon.read_dft(c)
        ref = omon.dft(c)
        peak = max(float(np.max(np.abs(omon.dft(q)))) for q in FIELDS)
        assert float(np.max(np.abs(got - ref))) / peak < 1e-5, c
    faces = mon.fetch_dft_fields()
    assert set(faces) == set(FIELDS) and mon.Ex_dft.shape == tuple(case["shape"])
    assert mon.Ex_dft.dtype == np.complex64


def test_farfield_api_errors_and_shapes(cases):
    case = dict(cases["vacuum_farfield"], n_steps=40)
    _, mon = run_farfield(case)
    eh = mon.get_farfields(np.array([[0, 0, 10.0], [10.0, 0, 0]], np.float32))
    assert eh.shape == (2, 6) and eh.dtype == np.complex128
    assert mon.get_farfields(np.array([0, 0, 10.0])).shape == (1, 6)
    with pytest.raises(ValueError):
        mon.get_farfields(np.ones((3, 2)))
    with pytest.raises(ValueError):
        mon.farfield_polar_xz(distance_m=0.0)


def test_multifrequency_monitor_equals_single_frequency_monitors(cases):
    case = dict(cases["dielectric_sphere_farfield"], n_steps=200)
    freqs = [4e9, 5e9, 6.5e9]
    _, multi = run_farfield(case, freqs=freqs)
    for q, f in enumerate(freqs):
        _, single = run_farfield(case, mon_freq=f)
        for c in FIELDS:
            np.testing.assert_array_equal(multi.read_dft(c, q), single.read_dft(c), err_msg=c)
        pts = [(0, 0, 3.0), (3.0, 0, 0), (1.0, 1.0, 1.0)]
        np.testing.assert_array_equal(multi.get_farfields(pts, q), single.get_farfields(pts))


def test_accumulator_modes_against_fp64(cases):
    """Over a 2000-step run, Kahan-compensated FP32 sums track the FP64 ones
    more closely than plain per-step FP32 sums; the K-buffered matrix-core
    sums (K = 32, the FP32 default) sit between them, and every mode stays
    within 1e-5 of FP64 (peak-normalised)."""
    case = dict(cases["dielectric_sphere_farfield"], n_steps=2000)
    runs = {"fp32": ("fp32", 0), "kahan": ("kahan", 0), "fp64": ("fp64", 0),
            "buffered": ("fp32", 32), "buffered8": ("fp32", 8)}  # fmt: skip
    res = {k: run_farfield(case, accum=a, buffer_steps=b)[1] for k, (a, b) in runs.items()}
    err = {}
    for m in ("fp32", "kahan", "buffered", "buffered8"):
        err[m] = max(fo.peak_normalised_error(res[m].read_dft(c), res["fp64"].read_dft(c))
                     for c in FIELDS)  # fmt: skip
    print("accumulator error vs fp64:", err, flush=True)
    assert res["buffered"].buffer_steps == 32 and res["fp32"].buffer_steps == 0
    assert err["kahan"] < err["fp32"]
    assert err["kahan"] < 1e-6
    assert err["buffered"] < 1e-5 and err["buffered8"] < 1e-5
    assert err["buffered"] <= err["fp32"]


def test_relative_change_matches_host_formula(cases):
    case = dict(cases["vacuum_farfield"], n_steps=100)
    sim, mon = run_farfield(case)
    assert mon.dft_relative_change() == 1.0
    mon.snapshot_dft()
    prev = {c: mon.read_dft(c).astype(np.complex128) for c in FIELDS}
    sim.run(20)
    cur = {c: mon.read_dft(c).astype(np.complex128) for c in FIELDS}
    num = sum(float(np.sum(np.abs(cur[c] - prev[c]) ** 2)) for c in FIELDS)
    den = sum(float(np.sum(np.abs(cur[c]) ** 2)) for c in FIELDS)
    assert mon.dft_relative_change() == pytest.approx(np.sqrt(num / den), rel=1e-6)


def test_full_size_linearity_and_fused_vs_split():
    """256^3 (npml 12): doubling the source doubles every field bit-exactly,
    and the fused sweep tracks the split sweeps."""
    shape, npml = (256, 256, 256), 12
    eps = np.ones(shape, np.float32)
    eps[100:160, 90:170, 60:200] = 3.0
    res = {}
    for key, amp, fused in (("a", 1.0, True), ("b", 2.0, True), ("s", 1.0, False)):
        print(f"[256^3] run {key}", flush=True)
        sim = hip().HIPFDTD(shape, 1e-3, npml=npml, fused=fused)
        sim.set_epsilon(eps)
        sim.add_source(lambda f, a=amp: f.add_source_Jx(
            128, a * fo.gaussian_sine(f.t, 1.5e10, 7.5e9), 112, 144, 112, 144, rim_taper=True))  # fmt: skip
        sim.run(60)
        res[key] = sim.fields()
        sim.close()
    # exact except where values pass through the FP32 denormal range (|x| < 1.2e-38)
    doubled = {n: 2.0 * res["a"][n] for n in FIELDS}
    assert group_error(res["b"], doubled) < 1e-20
    assert group_error(res["a"], res["s"]) < 1e-6
