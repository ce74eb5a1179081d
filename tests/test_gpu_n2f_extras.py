"""Near-to-far monitor options the fixtures do not cover, on the GPU.

* ``fetch_dft_fields`` (reference ``monitors.py:757-804``): the sparse
  volumes it scatters must hold, at every face cell, the face-pack value of
  the last face (in packing order x0, x1, y0, y1, z0, z1) that contains the
  cell, and zero elsewhere.  The expected volumes are built here sample by
  sample from the in-face index formulas (``monitors.py:396-412``) and the
  oracle's face DFT.
* ``get_farfields(l_sign=+1)``: the NumPy monitor's ``η0 N⊥ + r̂×L`` form
  (``monitors.py:171, 378``) against the oracle's ``l_sign=+1`` on the same
  runs.
"""

from __future__ import annotations

import numpy as np
import pytest
from test_oracle_fixtures import run_farfield_case
from test_gpu_parity import run_farfield

from oracle import fdtd_oracle as fo
from oracle.n2f_oracle import face_samples

pytestmark = pytest.mark.gpu

FIELDS = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")


def _scatter_by_sample(shape, box, packs, dtype):
    """Sparse volume from a face pack, one packed sample at a time."""
    vol = np.zeros(shape, dtype=dtype)
    face, I, J, K = face_samples(box)
    for q in np.argsort(face, kind="stable"):  # later faces overwrite shared edges
        vol[I[q], J[q], K[q]] = packs[q]
    return vol


@pytest.mark.parametrize("name", ["vacuum_farfield", "dielectric_sphere_farfield"])
def test_fetch_dft_fields_volumes(cases, name):
    case = cases[name]
    print(f"[fetch-dft] {name}", flush=True)
    sim, mon = run_farfield(case)
    _, omon = run_farfield_case(case)
    faces = mon.fetch_dft_fields()
    shape = tuple(case["shape"])
    peak = max(float(np.max(np.abs(omon.dft(c)))) for c in FIELDS)
    mask = _scatter_by_sample(shape, mon._box, np.ones(mon.n_face_samples), np.float64) > 0
    for c in FIELDS:
        vol = getattr(mon, f"{c}_dft")
        assert vol.shape == shape and vol.dtype == np.complex64
        np.testing.assert_array_equal(faces[c], mon.read_dft(c))
        # bitwise: the volume holds the library's face packs
        want = _scatter_by_sample(shape, mon._box, faces[c], np.complex64)
        np.testing.assert_array_equal(vol, want, err_msg=c)
        assert not np.any(vol[~mask]), c
        # and the packs are the oracle's face DFT
        ref = _scatter_by_sample(shape, mon._box, omon.dft(c), np.complex128)
        assert float(np.max(np.abs(vol - ref))) / peak < 1e-5, c


@pytest.mark.parametrize("name", ["vacuum_farfield", "dielectric_sphere_farfield"])
def test_legacy_l_sign_matches_oracle(cases, name):
    case = cases[name]
    print(f"[l-sign] {name}", flush=True)
    _, mon = run_farfield(case)
    _, omon = run_farfield_case(case)
    R = case["obs_r"]
    rad = np.deg2rad(np.linspace(-180.0, 180.0, 37))
    pts = np.column_stack([R * np.sin(rad), 0.3 * R * np.cos(rad), R * np.cos(rad)])
    kw = dict(k_dtype=np.float32, ang_dtype=np.float32)
    for sign in (+1, -1):
        got = mon.get_farfields(pts, l_sign=sign)
        ref = omon.farfields(pts.astype(np.float32), l_sign=float(sign), **kw)
        for grp in (slice(0, 3), slice(3, 6)):
            assert fo.peak_normalised_error(got[:, grp], ref[:, grp]) < 1e-4, (sign, grp)
    # the two forms differ, so the switch is live
    a, b = mon.get_farfields(pts, l_sign=+1), mon.get_farfields(pts, l_sign=-1)
    assert fo.peak_normalised_error(a, b) > 0.1
