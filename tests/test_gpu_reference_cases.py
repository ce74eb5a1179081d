"""The reference's own engine parity cases, restated against the oracle.

Each case keeps the reference test's grid, materials, source, step count and
tolerance, with the oracle's two CPU twins standing in for the reference's
NumPy engines: ``OracleVolumeFDTD`` (full-volume psi, as ``NumPyFDTD``) and
``OracleFDTD`` (face-stripe psi, as ``NumPyFDTD_FaceCPML``).

* tests/test_solver.py:32-46,47-111 -- 30^3 grid, dl 2 mm, npml 6, an eps_r 4
  sphere of radius 6 at the centre, a full Ex sheet sin(2 pi 5 GHz t) at
  k = nz - npml - 2, 50 steps, every field within 1e-4 of both twins and
  nonzero;
* tests/test_unit_engine.py:322-338 -- 24^3 vacuum, dl 2 mm, npml 4,
  4 GHz sheet, 20 steps, within 2e-4 (Ey and Hx stay zero by symmetry);
* tests/test_unit_engine.py:438-474 -- 20^3, dl 1 mm, npml 3, an eps_r 2.5
  patch in the source plane, add_source_Jx(10, 0.4) on the interior sheet
  with rim taper 0.8 and renormalisation, 15 steps, within 2e-4;
* tests/test_unit_engine.py:135-145 -- the Courant time step
  0.99 dl / (c sqrt 3) to 16 places and the INFO progress log of run().

Every step path of the engine runs each case; the two-steps-per-launch
path ('pair') advances through run().
"""

from __future__ import annotations

import logging

import numpy as np
import pytest

from oracle import fdtd_oracle as fo

pytestmark = [pytest.mark.gpu, pytest.mark.filterwarnings("ignore:add_source_Ex is a legacy")]

FIELDS = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")
PATHS = ("split", "fused", "march", "pair")


def hip():
    import hipfdtd

    return hipfdtd


def _advance(sim, path, steps):
    if path == "pair":
        sim.run(steps)
    else:
        for _ in range(steps):
            sim.step()


def _compare(sim, twins, tol, label, nonzero=False):
    got = sim.fields()
    for name, twin in twins.items():
        want = twin.fields()
        for f in FIELDS:
            if nonzero:
                assert bool(np.any(want[f] != 0)), f"{label} {name} {f} is zero"
            diff = float(np.max(np.abs(got[f].astype(np.float64) - want[f].astype(np.float64))))
            assert diff < tol, f"{label} vs {name}: {f} differs by {diff:.3e}"


def _sphere_eps(shape, ctr=15, rad=6):
    i, j, k = np.indices(shape)
    eps = np.ones(shape, np.float32)
    eps[(i - ctr) ** 2 + (j - ctr) ** 2 + (k - ctr) ** 2 < rad**2] = 4.0
    return eps


@pytest.mark.parametrize("path", PATHS)
def test_sphere_fields_match_both_numpy_twins(path):
    """tests/test_solver.py:47-111: 30^3, eps_r 4 sphere, 50 steps, 1e-4."""
    shape, dl, npml, freq = (30, 30, 30), 2e-3, 6, 5e9
    print(f"[ref-case] sphere 30^3 path={path}", flush=True)
    eps = _sphere_eps(shape)
    z_src = shape[2] - npml - 2
    twins = {"volume-psi": fo.OracleVolumeFDTD(shape, dl, npml=npml),
             "face-psi": fo.OracleFDTD(shape, dl, npml=npml)}  # fmt: skip
    sim = hip().HIPFDTD(shape, dl, npml=npml, path=path)
    for s in (sim, *twins.values()):
        s.set_epsilon(eps)
        s.add_source(lambda f: f.add_source_Ex(z_src, np.sin(2 * np.pi * freq * f.t)))
    _advance(sim, path, 50)
    for t in twins.values():
        t.run(50)
    assert sim.step_num == 50
    _compare(sim, twins, 1e-4, f"sphere {path}", nonzero=True)
    sim.close()


@pytest.mark.parametrize("path", PATHS)
def test_vacuum_sheet_matches_numpy_twin(path):
    """tests/test_unit_engine.py:322-338: 24^3 vacuum, 20 steps, 2e-4."""
    shape, dl, npml, freq = (24, 24, 24), 2e-3, 4, 4e9
    print(f"[ref-case] vacuum 24^3 path={path}", flush=True)
    z = shape[2] - npml - 2
    twins = {"volume-psi": fo.OracleVolumeFDTD(shape, dl, npml=npml)}
    sim = hip().HIPFDTD(shape, dl, npml=npml, path=path)
    for s in (sim, *twins.values()):
        s.add_source(lambda f: f.add_source_Ex(z, np.sin(2 * np.pi * freq * f.t)))
    _advance(sim, path, 20)
    twins["volume-psi"].run(20)
    _compare(sim, twins, 2e-4, f"vacuum {path}")
    sim.close()


@pytest.mark.parametrize("path", PATHS)
def test_tapered_renormalised_jx_on_dielectric_patch(path):
    """tests/test_unit_engine.py:438-474: rim-tapered, renormalised Jx sheet
    over an eps_r 2.5 patch, 15 steps, 2e-4."""
    shape, dl, npml, z, jx = (20, 20, 20), 1e-3, 3, 10, 0.4
    print(f"[ref-case] Jx rim 20^3 path={path}", flush=True)
    p = npml
    eps = np.ones(shape, np.float32)
    eps[8:12, 8:12, z] = 2.5

    def src(f):
        f.add_source_Jx(z, jx, i0=p, i1=shape[0] - p, j0=p, j1=shape[1] - p,
                        rim_taper=True, rim_edge=0.8, rim_renorm=True)  # fmt: skip

    twins = {"volume-psi": fo.OracleVolumeFDTD(shape, dl, npml=npml)}
    sim = hip().HIPFDTD(shape, dl, npml=npml, path=path)
    for s in (sim, *twins.values()):
        s.set_epsilon(eps)
        s.add_source(src)
    _advance(sim, path, 15)
    twins["volume-psi"].run(15)
    _compare(sim, twins, 2e-4, f"Jx rim {path}")
    sim.close()


def test_courant_dt_and_progress_log(caplog):
    """tests/test_unit_engine.py:135-145."""
    sim = hip().HIPFDTD((16, 16, 16), 1e-3, npml=4)
    assert sim.dt == pytest.approx(0.99 * 1e-3 / (299_792_458.0 * np.sqrt(3.0)), abs=1e-16)
    with caplog.at_level(logging.INFO, logger="hipfdtd.engine"):
        sim.run(3, progress_every=1)
    out = "\n".join(r.getMessage() for r in caplog.records if r.name == "hipfdtd.engine")
    assert "step 0/3" in out
    assert sim.step_num == 3
    sim.close()


def test_grid_beyond_device_memory_raises_memory_error():
    """A grid larger than the device's usable memory fails in the
    constructor, before any allocation, naming the usable budget
    (tests/test_unit_engine.py:244-257 checks the same on a mocked device)."""
    with pytest.raises(MemoryError) as err:
        hip().HIPFDTD((4096, 4096, 4096), 1e-3, npml=20)
    assert "usable" in str(err.value).lower()
