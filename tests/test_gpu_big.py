"""The march sweep on FP32 slabs of 4 GB and more.

Past 4 GB the slab's 32-bit buffer offsets no longer reach every volume, so
the sweep addresses each volume through a resource of its own (volumes
below 2 GB, about 810^3 cells).  On small grids that addressing can be
forced: it must give the slab-addressed sweep's results bit for bit, for
every medium kind and tiling, with sources in the sweep and a Huygens-box
monitor.  At 576^3 (slab 4.3 GB) the engine takes it by itself and must
track the split sweeps (1e-6 of the field-group peak, as the one-launch
sweeps do elsewhere).  The reference publishes 600^3 and 750^3 runs
(README.md:189-190).
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import fdtd_oracle as fo

pytestmark = pytest.mark.gpu

FIELDS = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")


def hip():
    import hipfdtd

    return hipfdtd


def _medium(shape, kind, seed=3):
    rng = np.random.default_rng(seed)
    if kind == "indexed":
        eps = np.ones(shape, np.float32)
        for v in (2.0, 3.5, 5.0):
            lo = [int(rng.integers(0, n // 2)) for n in shape]
            eps[lo[0]:lo[0] + shape[0] // 3, lo[1]:lo[1] + shape[1] // 3, lo[2]:lo[2] + 20] = v
        return eps, None
    eps = rng.uniform(1.0, 6.0, size=shape).astype(np.float32)
    sig = rng.uniform(0.0, 2.0, size=shape).astype(np.float32) if kind == "lossy" else None
    return eps, sig


def _drive(sim):
    nx, ny, nz = sim.Nx, sim.Ny, sim.Nz

    def src(f):
        a = fo.gaussian_sine(f.t, 3e10, 1.5e10)
        f.add_source_Jx(nz // 2, a, 2, nx - 3, 3, ny - 2, rim_taper=True, rim_edge=0.7)
        f.add_source_J("z", 0.5 * a, (nx // 3, nx // 3 + 1, ny // 2, ny // 2 + 1, 5, 6))

    sim.add_source(src)


@pytest.mark.parametrize("kind", ["lossless", "lossy", "indexed"])
@pytest.mark.parametrize("rows", [1, 2])
def test_per_volume_addressing_matches_slab_addressing(kind, rows):
    shape, dl = (45, 35, 100), 1e-3
    print(f"[per-volume] kind={kind} rows={rows} shape={shape}", flush=True)
    eps, sig = _medium(shape, kind)
    out = []
    for per_volume in (False, True):
        sim = hip().HIPFDTD(shape, dl, npml=6, path="march")
        sim.set_march_tiling(rows, 0, 0, 1)
        sim.set_march_addressing(per_volume)
        sim.set_epsilon(eps, sig)
        assert (sim.material_count() > 0) == (kind == "indexed")
        _drive(sim)
        mon = hip().HIPNear2FarMonitor(sim, [s * dl / 2 for s in shape],
                                       [(s - 16) * dl for s in shape], freqs=[2.5e10, 3e10])  # fmt: skip
        sim.run(14)
        out.append((sim.fields(), [mon.read_dft(c, 1) for c in FIELDS]))
    (fa, da), (fb, db) = out
    for n in FIELDS:
        np.testing.assert_array_equal(fa[n], fb[n], err_msg=n)
    for a, b in zip(da, db):
        np.testing.assert_array_equal(a, b)


def test_576cube_takes_per_volume_march_and_tracks_split():
    shape, npml = (576, 576, 576), 12
    print(f"[big] {shape}: march (per-volume addressing) against split, 4 steps", flush=True)
    eps = np.ones(shape, np.float32)
    eps[200:380, 150:300, 250:420] = 3.0
    res = {}
    for path in ("march", "split"):
        sim = hip().HIPFDTD(shape, 1e-3, npml=npml, path=path)
        assert sim.path == path
        sim.set_epsilon(eps)
        sim.add_source(lambda f: f.add_source_Jx(
            288, fo.gaussian_sine(f.t, 1.5e10, 7.5e9), 256, 320, 256, 320, rim_taper=True))  # fmt: skip
        sim.run(4)
        res[path] = sim.fields()
        sim.close()
    a, b = res["march"], res["split"]
    for grp in (FIELDS[:3], FIELDS[3:]):
        peak = max(float(np.max(np.abs(b[n]))) for n in grp)
        assert peak > 0
        for n in grp:
            err = float(np.max(np.abs(a[n].astype(np.float64) - b[n]))) / peak
            assert err < 1e-6, (n, err)
