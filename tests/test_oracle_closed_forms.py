"""Closed forms and invariants that pin parts of the oracle the fixtures leave open.

Mirrors the assertions of the reference's test_unit_cpml.py:24-50,
test_materials_lossy.py:17-56 and test_unit_engine.py:54-72, 149-243.
"""

from __future__ import annotations

import math

import numpy as np
import pytest

from oracle import fdtd_oracle as fo
from oracle.n2f_oracle import OracleNear2Far, face_layout, snap_box


def test_sigma_max_taflove():
    assert fo.cpml_sigma_max(1e-3) == pytest.approx(0.8 * 4 / (fo.ETA0 * 1e-3), rel=1e-14)


def test_cpml_kappa_grading_and_stagger():
    dl, dt, npml, n = 1e-3, 1e-12, 8, 40
    be, ce, ke = fo.cpml_profile(n, npml, dl, dt, 0.0)
    bh, ch, kh = fo.cpml_profile(n, npml, dl, dt, 0.5)
    assert ke[0] == pytest.approx(fo.CPML_KAPPA_MAX, abs=1e-5)  # outer wall, xi = 1
    assert ke[n - 1] == pytest.approx(fo.CPML_KAPPA_MAX, abs=1e-5)
    assert ke[n // 2] == 1.0 and be[n // 2] == 1.0 and ce[n // 2] == 0.0
    assert not np.allclose(be[:npml], bh[:npml])
    assert not np.allclose(ke[:npml], kh[:npml])
    # H node on the hi side: xi = (i + 1 - 1/2)/npml; check one entry by hand
    i = 3
    xi = (i + 0.5) / npml
    sig = fo.cpml_sigma_max(dl) * xi**3
    kap = 1 + 14 * xi**3
    alp = 0.05 / fo.ETA0 * (1 - xi)
    b = np.float32(math.exp(-(sig / kap + alp) * dt / fo.EPS0))
    assert bh[n - npml + i] == b
    assert ch[n - npml + i] == np.float32(sig / kap * (float(b) - 1) / (sig + kap * alp) / dl)


def test_materials_zero_sigma_is_lossless():
    eps = np.array([[[1.0, 2.0, 3.0]]])
    (cax, cay, caz), (cbx, cby, cbz) = fo.edge_ca_cb(eps, np.zeros_like(eps), 1e-13, np.float64)
    ce = fo.edge_ce(eps, 1e-13, np.float64)
    for a in (cax, cay, caz):
        np.testing.assert_array_equal(a, np.ones_like(a))
    for b, c in zip((cbx, cby, cbz), ce):
        np.testing.assert_allclose(b, c, rtol=0, atol=0)


def test_materials_hand_formula_and_bounds():
    eps_r, s, dt = 2.5, 15.0, 3e-13
    (ca, _, _), (cb, _, _) = fo.edge_ca_cb(np.full((3, 3, 3), eps_r), np.full((3, 3, 3), s), dt,
                                            np.float64)  # fmt: skip
    loss = s * dt / (2 * fo.EPS0 * eps_r)
    assert ca[1, 1, 1] == pytest.approx((1 - loss) / (1 + loss), rel=1e-12)
    assert cb[1, 1, 1] == pytest.approx(dt / (fo.EPS0 * eps_r) / (1 + loss), rel=1e-12)
    for sv in (100.0, 1e3, 1e6):
        (ca, _, _), _ = fo.edge_ca_cb(np.ones((2, 2, 2)), np.full((2, 2, 2), sv), 1e-12, np.float64)
        assert np.all(ca > -1) and np.all(ca <= 1)


def test_edge_average_last_plane_keeps_cell_value():
    eps = np.ones((12, 12, 12), np.float32)
    eps[6, 6, 6] = 4.0
    dt = fo.OracleFDTD((12, 12, 12), 1e-3, npml=2).dt
    cex, cey, cez = fo.edge_ce(eps, dt)
    assert cex[6, 6, 6] == np.float32(dt / (fo.EPS0 * 2.5))
    assert cex[5, 6, 6] == np.float32(dt / (fo.EPS0 * 2.5))
    assert cex[0, 0, 0] == np.float32(dt / fo.EPS0)
    eps2 = np.ones((4, 4, 4))
    eps2[-1] = 3.0
    assert fo.edge_average(eps2)[0][-1, 0, 0] == 3.0


def test_face_psi_matches_volume_psi():
    """Face-stripe psi (OpenCL layout) equals a full-volume psi restatement."""
    shape, dl, npml = (24, 24, 24), 1e-3, 4
    a = fo.OracleFDTD(shape, dl, npml=npml)
    b = fo.OracleVolumeFDTD(shape, dl, npml=npml)
    z = shape[2] // 2

    def src(f):
        f.Ex[npml:-npml, npml:-npml, z] += np.float32(np.sin(2 * np.pi * 5e9 * f.t))

    a.add_source(src)
    b.add_source(src)
    for _ in range(40):
        a.step()
        b.step()
    # normalised by the peak of the field group: Hx is pure round-off here
    for group in (("Ex", "Ey", "Ez"), ("Hx", "Hy", "Hz")):
        peak = max(float(np.max(np.abs(getattr(b, n)))) for n in group)
        for name in group:
            diff = float(np.max(np.abs(getattr(a, name) - getattr(b, name))))
            assert diff / peak < 1e-5, name


def test_jx_injection_closed_form():
    shape, dl, npml, z, J = (16, 16, 16), 1e-3, 2, 8, 0.25
    sim = fo.OracleFDTD(shape, dl, npml=npml)
    eps = np.ones(shape, np.float32)
    eps[8, 8, z] = 4.0
    sim.set_epsilon(eps)
    p = npml
    sim.add_source(lambda f: f.add_source_Jx(z, J, i0=p, i1=16 - p, j0=p, j1=16 - p))
    sim.step()
    assert sim.Ex[8, 8, z] == pytest.approx(-sim.dt / (fo.EPS0 * 2.5) * J, rel=1e-6)
    assert sim.Ex[p, p, z] == pytest.approx(-sim.dt / fo.EPS0 * J, rel=1e-6)
    assert sim.Ex[0, 0, z] == 0.0


def test_jx_rim_weights():
    shape, npml, z, ew = (16, 16, 16), 2, 8, 0.5
    p, i1, j1 = npml, 16 - npml, 16 - npml
    sim = fo.OracleFDTD(shape, 1e-3, npml=npml)
    sim.add_source(lambda f: f.add_source_Jx(z, 1.0, i0=p, i1=i1, j0=p, j1=j1, rim_taper=True,
                                             rim_edge=ew, rim_renorm=False))  # fmt: skip
    sim.step()
    base = -sim.dt / fo.EPS0
    assert sim.Ex[p + 2, p + 2, z] == pytest.approx(base, rel=1e-6)
    assert sim.Ex[p, p + 2, z] == pytest.approx(ew * base, rel=1e-6)
    assert sim.Ex[p + 2, p, z] == pytest.approx(ew * base, rel=1e-6)
    assert sim.Ex[p, p, z] == pytest.approx(ew * ew * base, rel=1e-6)
    assert sim.Ex[i1 - 1, j1 - 1, z] == pytest.approx(ew * ew * base, rel=1e-6)


def test_rim_renorm_preserves_total_current():
    shape, z = (20, 20, 20), 10
    sim = fo.OracleFDTD(shape, 1e-3, npml=2)
    sim.add_source(lambda f: f.add_source_Jx(z, 1.0, 3, 15, 4, 17, rim_taper=True))
    sim.step()
    total = float(np.sum(sim.Ex[:, :, z], dtype=np.float64))
    assert total == pytest.approx(-sim.dt / fo.EPS0 * 12 * 13, rel=1e-5)


def test_npml0_closed_box_energy_bounded():
    """PEC-like box (npml=0): energy after the source stops stays bounded."""
    sim = fo.OracleFDTD((20, 20, 20), 1e-3, npml=0, dtype=np.float64)
    sim.add_source(lambda f: f.add_source_Ex(10, fo.gaussian_sine(f.t, 3e10, 1e10), 8, 12, 8, 12))
    energies = []
    for s in range(200):
        sim.step()
        if s >= 120:
            e = sum(float(np.sum(v.astype(np.float64) ** 2)) for v in (sim.Ex, sim.Ey, sim.Ez))
            h = sum(float(np.sum(v.astype(np.float64) ** 2)) for v in (sim.Hx, sim.Hy, sim.Hz))
            energies.append(fo.EPS0 * e + fo.MU0 * h)
    assert max(energies) <= 1.15 * energies[0]


def test_box_snapping_and_layout():
    box = snap_box((30, 30, 30), 2e-3, (30e-3, 30e-3, 30e-3), (20e-3, 20e-3, 20e-3))
    assert box == (10, 20, 10, 20, 10, 20)
    counts, offs, n = face_layout(box)
    assert counts == (121,) * 6 and offs == (0, 121, 242, 363, 484, 605) and n == 726
    # round-half-even: 2.5 -> 2
    assert snap_box((30, 30, 30), 1.0, (5.0, 5.0, 5.0), (5.0, 5.0, 5.0))[0] == 2


def test_phase_recurrence_matches_direct():
    sim = fo.OracleFDTD((20, 20, 20), 1e-3, npml=3)
    mon = OracleNear2Far(sim, (0.01, 0.01, 0.01), (0.008, 0.008, 0.008), 3e9)
    sim.run(5)
    direct = np.exp(1j * mon.omegas * sim.t) * sim.dt
    np.testing.assert_allclose(mon._phase, direct, rtol=1e-12)


def test_box_current_source_closed_form():
    """add_source_J on Ez and Ey boxes: ΔE = -dt/(ε0 εe) J · w."""
    sim = fo.OracleFDTD((10, 11, 12), 1e-3, npml=2)
    eps = np.ones((10, 11, 12), np.float32)
    eps[4, 5, 6] = 2.0
    sim.set_epsilon(eps)
    sim.add_source(lambda f: f.add_source_J("z", 2.0, (4, 5, 5, 6, 6, 7)))
    sim.add_source(lambda f: f.add_source_J("y", 1.0, (2, 5, 3, 4, 1, 4), rim_taper=5,
                                            rim_edge=0.5))  # fmt: skip
    sim.step()
    assert sim.Ez[4, 5, 6] == pytest.approx(-sim.dt / (fo.EPS0 * 1.5) * 2.0, rel=1e-6)
    base = -sim.dt / fo.EPS0
    assert sim.Ey[3, 3, 2] == pytest.approx(base, rel=1e-6)          # interior of the box
    assert sim.Ey[2, 3, 2] == pytest.approx(0.5 * base, rel=1e-6)    # i edge
    assert sim.Ey[2, 3, 1] == pytest.approx(0.25 * base, rel=1e-6)   # i and k edges


def test_l_sign_acts_on_the_magnetic_current_term_only():
    """E = pref (η0 N⊥ + l r̂×L): with only H sampled on the faces L = 0 and
    both signs agree; with only E sampled N = 0 and they are opposite.  The
    POCL fixtures pin the l = -1 form; this pins what l = +1 changes."""
    sim = fo.OracleFDTD((20, 20, 20), 1e-3, npml=3)
    mon = OracleNear2Far(sim, (0.01, 0.01, 0.01), (0.008, 0.008, 0.008), 3e9)
    rng = np.random.default_rng(4)
    pts = rng.normal(size=(7, 3))
    for c in mon.acc:
        mon.acc[c][:] = rng.normal(size=mon.acc[c].shape)
    full = {c: v.copy() for c, v in mon.acc.items()}
    for keep, sign in (("H", 1.0), ("E", -1.0)):
        for c in mon.acc:
            mon.acc[c][:] = full[c] if c[0] == keep else 0.0
        a = mon.farfields(pts, l_sign=+1.0)
        b = mon.farfields(pts, l_sign=-1.0)
        assert np.max(np.abs(a)) > 0
        np.testing.assert_allclose(a, sign * b, rtol=1e-13, atol=1e-13 * np.max(np.abs(a)))
