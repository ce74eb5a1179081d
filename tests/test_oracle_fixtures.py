"""The oracle against the reference's committed POCL FP32 outputs.

Plane fixtures pin the whole Yee + face-CPML + Jx-sheet path
(plane_cases.py:116-146); far-field fixtures pin the face DFT, the N/L
surface integral and the N,L -> E,H transform (cases.py:249-264).  A planted
mistake in the oracle must break the agreement.
"""

from __future__ import annotations

import numpy as np
import pytest
from conftest import golden, plane_eps, sphere_eps

from oracle import fdtd_oracle as fo
from oracle.n2f_oracle import OracleNear2Far


def run_plane_case(case, cls=fo.OracleFDTD, dtype=np.float32):
    nx, ny, nz = case["shape"]
    p = case["npml"]
    sim = cls(tuple(case["shape"]), case["dl"], npml=p, dtype=dtype)
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
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_oracle_matches_plane_fixtures(cases, name, dtype):
    planes = run_plane_case(cases[name], dtype=dtype)
    for step, got in planes.items():
        ref = golden(name, f"ocl_ex_step{step:04d}.npy")
        err = fo.peak_normalised_error(got, ref)
        # FP32 POCL vs this restatement: observed <= 4e-7; bound 2e-6
        assert err < 2e-6, (name, step, err)


class _WrongStagger(fo.OracleFDTD):
    """Planted mistake: H uses the E-node CPML profile."""

    def _build_cpml(self):
        super()._build_cpml()
        self.prof_h = self.prof_e


class _WrongSourceSign(fo.OracleFDTD):
    """Planted mistake: Ex += +ce*J instead of -ce*J."""

    def add_source_Jx(self, z_src, Jx, *a, **k):
        return super().add_source_Jx(z_src, -Jx, *a, **k)


class _NoPsiZ(fo.OracleFDTD):
    """Planted mistake: the z-stripe psi recursion is dropped."""

    def _psi_update(self, key, axis, prof, d):
        vol = super()._psi_update(key, axis, prof, d)
        return vol * 0 if axis == 2 else vol


@pytest.mark.parametrize("cls", [_WrongStagger, _WrongSourceSign, _NoPsiZ])
def test_planted_oracle_mistakes_fail_fixtures(cases, cls):
    planes = run_plane_case(cases["vacuum_sheet"], cls=cls)
    worst = max(
        fo.peak_normalised_error(got, golden("vacuum_sheet", f"ocl_ex_step{s:04d}.npy"))
        for s, got in planes.items()
    )
    assert worst > 1e-3, worst


def run_farfield_case(case, accum=np.float32, dtype=np.float32):
    sim = fo.OracleFDTD(tuple(case["shape"]), case["dl"], npml=case["npml"], dtype=dtype)
    if case["sphere_eps"]:
        sim.set_epsilon(sphere_eps(case))
    i0, i1, j0, j1, kk = case["patch"]
    sim.add_source(
        lambda f: f.add_source_Ex(kk, fo.gaussian_sine(f.t, case["freq"], case["fwidth"]),
                                  i0, i1, j0, j1)
    )  # fmt: skip
    mon = OracleNear2Far(sim, case["monitor_center"], case["monitor_size"], case["freq"],
                         accum=accum)  # fmt: skip
    sim.run(case["n_steps"])
    return sim, mon


@pytest.mark.parametrize("name", ["vacuum_farfield", "dielectric_sphere_farfield"])
def test_oracle_matches_farfield_fixtures(cases, name):
    case = cases[name]
    _, mon = run_farfield_case(case)
    assert mon.box == (10, 20, 10, 20, 10, 20)
    R = case["obs_r"]
    ref_db = golden(name, "ocl_S_db.npy")
    ref_z = golden(name, "ocl_eh_plus_z.npy")
    ref_x = golden(name, "ocl_eh_plus_x.npy")
    # reference rounding: k in the field dtype, k*r in the accumulator dtype
    kw = dict(k_dtype=np.float32, ang_dtype=np.float32)
    ang, db = mon.farfield_polar_xz(R, case["n_angles"], **kw)
    np.testing.assert_allclose(ang, golden(name, "angles_deg.npy"), atol=1e-12)
    assert np.max(np.abs(db - ref_db)) < 1e-4
    ez = mon.farfields([(0, 0, R)], **kw)[0]
    ex = mon.farfields([(R, 0, 0)], **kw)[0]
    assert fo.peak_normalised_error(ez, ref_z) < 5e-6
    assert fo.peak_normalised_error(ex, ref_x) < 5e-6
    # the NumPy-monitor sign (+ r x L) does not reproduce the fixtures
    ez_plus = mon.farfields([(0, 0, R)], l_sign=+1.0, **kw)[0]
    assert fo.peak_normalised_error(ez_plus, ref_z) > 0.1
    # FP64 k moves E/H by the k*r phase error (~2.6e-5) but not |S|
    ez64 = mon.farfields([(0, 0, R)])[0]
    assert 1e-6 < fo.peak_normalised_error(ez64, ref_z) < 1e-4
