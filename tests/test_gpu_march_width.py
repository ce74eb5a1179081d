"""March tiles of 32 columns of k (two rows of j per wavefront slot).

The tile shape only regroups the same per-cell updates, so every width must
give the fields of the fused sweep bit for bit: on the material index, on
the eps_r volume, on the ce volumes with loss, without PML, with sources
applied inside the sweep and through the fix-up kernel (more than four), on
grids whose extents are below, at and between multiples of the width, and
with the Huygens box's z faces recorded by the sweep.  The batched march
sweep is checked the same way against serial engines.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import fdtd_oracle as fo

pytestmark = pytest.mark.gpu

FIELDS = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")


def hip():
    import hipfdtd
    import hipfdtd.batch
    import hipfdtd.probes

    return hipfdtd


def _medium(kind, shape, seed):
    rng = np.random.default_rng(seed)
    sig = None
    if kind == "index":
        eps = np.ones(shape, np.float32)
        for q in range(5):
            lo = [int(rng.integers(0, max(1, n - 6))) for n in shape]
            eps[lo[0]:lo[0] + 9, lo[1]:lo[1] + 7, lo[2]:lo[2] + 11] = 1.5 + 0.5 * q
    else:
        eps = rng.uniform(1.0, 6.0, size=shape).astype(np.float32)
        if kind == "lossy":
            sig = rng.uniform(0.0, 2.0, size=shape).astype(np.float32)
    return eps, sig


def _drive(sim, shape, steps, nsrc):
    nx, ny, nz = shape

    def src(f):
        a = fo.gaussian_sine(f.t, 3e10, 1.5e10)
        f.add_source_Jx(nz // 2, a, 2, nx - 3, 3, ny - 2, rim_taper=True, rim_edge=0.7)
        for q in range(nsrc - 1):
            f.add_source_J("xyz"[q % 3], (0.4 - 0.07 * q) * a,
                           (nx // 3 + q, nx // 3 + q + 1, ny // 2 - q, ny // 2 - q + 1,
                            min(nz - 1, 5 + 3 * q), min(nz, 6 + 3 * q)))  # fmt: skip

    sim.add_source(src)
    sim.run(steps)


def _fields(shape, npml, kind, path, nsrc, steps=14, width=0, planes=0, run=0):
    eps, sig = _medium(kind, shape, 5)
    sim = hip().HIPFDTD(shape, 1e-3, npml=npml, path=path)
    if path == "march":
        sim.set_march_width(width)
        sim.set_march_tiling(0, planes, run, -2)
    sim.set_epsilon(eps, sig)
    if kind == "index":
        assert sim.material_count() > 0
    _drive(sim, shape, steps, nsrc)
    out = sim.fields()
    sim.close()
    return out


@pytest.mark.parametrize("width", [32])
@pytest.mark.parametrize("shape,npml,kind,nsrc,planes,run", [
    ((45, 35, 100), 6, "index", 3, 0, 0),
    ((45, 35, 100), 6, "eps", 4, 6, 3),
    ((33, 41, 70), 5, "lossy", 2, -1, 1),
    ((24, 30, 20), 0, "eps", 3, 4, 64),
    ((40, 19, 47), 4, "index", 6, 8, 16),
    ((30, 37, 64), 7, "lossy", 5, 1000, 5)])  # fmt: skip
def test_narrow_march_matches_fused(width, shape, npml, kind, nsrc, planes, run):
    print(f"[march-width] W={width} shape={shape} npml={npml} {kind} nsrc={nsrc} "
          f"planes={planes} run={run}", flush=True)  # fmt: skip
    want = _fields(shape, npml, kind, "fused", nsrc)
    got = _fields(shape, npml, kind, "march", nsrc, width=width, planes=planes, run=run)
    for n in FIELDS:
        np.testing.assert_array_equal(got[n], want[n], err_msg=n)


def test_narrow_march_against_oracle():
    shape, npml = (30, 27, 41), 5
    eps, _ = _medium("eps", shape, 5)
    ora = fo.OracleFDTD(shape, 1e-3, npml=npml)
    ora.set_epsilon(eps)
    _drive(ora, shape, 30, 3)
    for width in (32,):
        got = _fields(shape, npml, "eps", "march", 3, steps=30, width=width)
        assert fo.peak_normalised_error(got["Ex"], ora.Ex) < 1e-5
        assert fo.peak_normalised_error(got["Hy"], ora.Hy) < 1e-5


@pytest.mark.parametrize("width", [32])
def test_narrow_march_zface_record(width):
    """A buffered Huygens box (its z faces recorded by the sweep) on narrow
    tiles: DFT sums and far field equal those of 64-column tiles."""
    shape, npml = (40, 44, 90), 6
    res = []
    for w in (64, width):
        sim = hip().HIPFDTD(shape, 1e-3, npml=npml, path="march")
        sim.set_march_width(w)
        eps, _ = _medium("index", shape, 3)
        sim.set_epsilon(eps)
        lo, hi = (10, 9, 17), (30, 35, 66)  # z faces at k = 17 and 66
        c = [0.5 * (a + b) * sim.dl for a, b in zip(lo, hi)]
        s = [(b - a) * sim.dl for a, b in zip(lo, hi)]
        mon = hip().HIPNear2FarMonitor(sim, c, s, freqs=[2.5e10, 3e10])
        _drive(sim, shape, 40, 2)
        res.append(({c: mon.read_dft(c, 1) for c in FIELDS},
                    mon.get_farfields([(0.3, 0.2, 1.0), (0.0, 0.0, -1.0)], 0)))
        sim.close()
    for c in FIELDS:
        np.testing.assert_array_equal(res[1][0][c], res[0][0][c], err_msg=c)
    np.testing.assert_array_equal(res[1][1], res[0][1])


def test_march_width_rejects_bad_values():
    sim = hip().HIPFDTD((16, 16, 16), 1e-3, npml=2)
    for bad in (48, 16):
        with pytest.raises(ValueError, match="width"):
            sim.set_march_width(bad)
    sim.close()


@pytest.mark.parametrize("width,rows", [(32, 1)])
def test_batch_narrow_march_matches_serial(width, rows):
    """Batched narrow march tiles against serial engines on 64-column march
    tiles: fields and probe DFTs bit for bit."""
    from test_gpu_batch import _march_batch_case, sphere_case

    rng = np.random.default_rng(12)
    shapes = [tuple(int(v) for v in rng.integers(18, 100, size=3)) for _ in range(5)]
    media = [sphere_case(rng, s) for s in shapes]
    print(f"[batch-width] W={width} shapes={shapes}", flush=True)
    serial, sims, dfts = _march_batch_case(shapes, media, 33, 8, rows, 24, 16, width=width)
    for q, sim in enumerate(sims):
        ref_f, ref_p = serial[q]
        got = sim.fields()
        for n in FIELDS:
            np.testing.assert_array_equal(got[n], ref_f[n], err_msg=f"grid {q} {n}")
        np.testing.assert_array_equal(dfts[q], ref_p, err_msg=f"grid {q} probe")
        sim.close()
