"""Material-indexed sweeps (lossless FP32 grids with few distinct eps_r).

An edge's ce depends only on the eps_r of its two cells, so on a grid with
at most 48 distinct values the march sweep and the split E sweep (single
engine and batched) read one byte of material per cell and take ce from a
table computed exactly as the ce volumes are.  Results must equal the
sweeps that read the ce volumes bit for bit: the fused sweep for the march
sweep, and the split sweeps of the same engine with the index turned off
(a ce volume rewritten with its own values) for the split sweeps.  Grids
with more values, lossy grids and grids whose ce volumes were written by
hand fall back to the volumes.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import fdtd_oracle as fo

pytestmark = pytest.mark.gpu

FIELDS = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")
DL = 1e-3


def hip():
    import hipfdtd

    return hipfdtd


def _boxes_eps(shape, nvals, seed, dtype=np.float32):
    """Vacuum with random boxes of nvals - 1 distinct eps_r values."""
    rng = np.random.default_rng(seed)
    vals = rng.uniform(1.5, 9.0, size=nvals - 1)
    eps = np.ones(shape, dtype=np.float64)
    for q in range(3 * nvals):
        lo = [int(rng.integers(0, n - 2)) for n in shape]
        hi = [min(n, a + int(rng.integers(2, max(3, n // 3)))) for a, n in zip(lo, shape)]
        eps[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]] = vals[q % (nvals - 1)]
    flat = eps.reshape(-1)  # every value on at least one cell, vacuum too
    for q in range(nvals - 1):
        flat[(q * 7919 + 13) % flat.size] = vals[q]
    flat[(nvals * 7919 + 13) % flat.size] = 1.0
    return eps.astype(dtype)


def _run(path, shape, npml, eps, nsrc, steps, poke=None, rows=0, planes=0, materials=None):
    """Fields after ``steps`` steps on ``path``; on the march path ``rows``
    rows per wavefront (0: the engine's own choice, one row on indexed
    grids) and ``planes`` planes per block (0: from the grid), and the
    material index must hold ``materials`` values (0: off)."""
    sim = hip().HIPFDTD(shape, DL, npml=npml, path=path)
    if path == "march":
        sim.set_march_tiling(rows, planes, 0, -2)
    sim.set_epsilon(eps)
    if poke is not None:
        poke(sim)
    if path in ("march", "split") and materials is not None:
        assert sim.material_count() == materials, (sim.material_count(), materials)
    nx, ny, nz = shape

    def src(f):
        a = fo.gaussian_sine(f.t, 3e10, 1.5e10)
        f.add_source_Jx(nz // 2, a, 2, nx - 3, 3, ny - 2, rim_taper=True, rim_edge=0.7)
        for q in range(nsrc - 1):
            f.add_source_J("xyz"[q % 3], (0.4 - 0.1 * q) * a,
                           (nx // 3 + q, nx // 3 + q + 1, ny // 2, ny // 2 + 1, 5 + q, 6 + q))  # fmt: skip

    sim.add_source(src)
    sim.run(steps)
    return sim.fields()


def _same(a, b):
    for n in FIELDS:
        np.testing.assert_array_equal(a[n], b[n], err_msg=n)


@pytest.mark.parametrize("rows", [0, 1, 2])
@pytest.mark.parametrize("shape,npml,nvals,nsrc,eps_dtype,indexed", [
    ((45, 35, 100), 6, 5, 3, np.float32, True),
    ((40, 52, 70), 8, 48, 4, np.float64, True),
    ((40, 52, 70), 8, 48, 2, np.float32, True),
    ((40, 52, 70), 8, 49, 2, np.float32, False),
    ((33, 40, 130), 0, 12, 6, np.float32, True)])  # fmt: skip
def test_indexed_march_matches_fused(shape, npml, nvals, nsrc, eps_dtype, indexed, rows):
    """FP32 engines; ``eps_dtype`` is the dtype of the εr array handed to
    set_epsilon (both are indexed when at most 48 values occur)."""
    print(f"[mat] shape={shape} npml={npml} nvals={nvals} nsrc={nsrc} "
          f"eps {np.dtype(eps_dtype).name} rows={rows}", flush=True)  # fmt: skip
    eps = _boxes_eps(shape, nvals, 3, eps_dtype)
    assert len(np.unique(eps)) == nvals
    want = nvals if indexed else 0
    _same(_run("fused", shape, npml, eps, nsrc, 16),
          _run("march", shape, npml, eps, nsrc, 16, rows=rows, materials=want))  # fmt: skip


@pytest.mark.parametrize("rows", [0, 1, 2])
@pytest.mark.parametrize("case", ["f32", "f64-exact", "f64-inexact", "poked", "npml0"])
def test_eps_volume_march_matches_fused(case, rows):
    """A continuous random medium (every cell its own eps_r, no index): the
    march sweep recomputes ce from the eps_r volume when the values are
    exact in FP32, and reads the ce volumes otherwise; both equal the fused
    sweep bit for bit."""
    shape, npml = ((37, 45, 70), 0) if case == "npml0" else ((37, 45, 70), 6)
    rng = np.random.default_rng(11)
    eps = rng.uniform(1.0, 7.0, size=shape)
    if case != "f64-inexact":
        eps = eps.astype(np.float32)
        if case == "f64-exact":
            eps = eps.astype(np.float64)
    want = "volumes" if case in ("f64-inexact", "poked") else "eps"
    print(f"[eps] {case} rows={rows} -> {want}", flush=True)

    def check(sim):
        assert sim.material_count() == 0
        assert sim.ce_source() == want, (sim.ce_source(), want)

    def poke(sim):
        if case == "poked":
            _index_off(sim)
        check(sim)

    _same(_run("fused", shape, npml, eps, 3, 16),
          _run("march", shape, npml, eps, 3, 16, poke, rows=rows, materials=0))  # fmt: skip


def test_ce_source_follows_the_medium():
    shape = (20, 24, 30)
    sim = hip().HIPFDTD(shape, DL, npml=4)
    assert sim.ce_source() == "index"  # the default medium: one material
    rng = np.random.default_rng(2)
    sim.set_epsilon(rng.uniform(1, 3, size=shape).astype(np.float32))
    assert sim.ce_source() == "eps"
    sim.set_epsilon(_boxes_eps(shape, 4, 1))
    assert sim.ce_source() == "index"
    sim.set_epsilon(rng.uniform(1, 3, size=shape).astype(np.float32),
                    np.full(shape, 0.1, np.float32))  # fmt: skip
    assert sim.ce_source() == "volumes"
    d = hip().HIPFDTD(shape, DL, npml=4, dtype=np.float64)
    d.set_epsilon(rng.uniform(1, 3, size=shape).astype(np.float32))
    assert d.ce_source() == "volumes"


def _index_off(sim):
    """Rewrite ce_x with its own values: the same coefficients, no index."""
    sim.write_volume("ce_x", sim.read_volume("ce_x"))


@pytest.mark.parametrize("shape,npml,nvals,nsrc,eps_dtype,indexed", [
    ((45, 35, 100), 6, 5, 3, np.float32, True),
    ((40, 52, 70), 8, 48, 4, np.float64, True),
    ((40, 52, 70), 8, 49, 2, np.float32, False),
    ((33, 40, 130), 0, 12, 6, np.float32, True),
    ((29, 31, 17), 4, 2, 2, np.float32, True)])  # fmt: skip
def test_indexed_split_matches_volume_split(shape, npml, nvals, nsrc, eps_dtype, indexed):
    print(f"[mat-split] shape={shape} npml={npml} nvals={nvals} nsrc={nsrc}", flush=True)
    eps = _boxes_eps(shape, nvals, 5, eps_dtype)
    assert len(np.unique(eps)) == nvals
    _same(_run("split", shape, npml, eps, nsrc, 16, materials=nvals if indexed else 0),
          _run("split", shape, npml, eps, nsrc, 16, _index_off, materials=0))  # fmt: skip


def test_indexed_batch_split_matches_volume_batch():
    """A batch of indexed and unindexed engines on the split path against the
    same batch with every index off, and a lone indexed engine's serial
    split run."""
    shapes = [(30, 26, 40), (22, 35, 19), (41, 24, 33)]
    print(f"[mat-batch] shapes={shapes}", flush=True)
    epss = [_boxes_eps(sh, 3 + q, 6 + q) for q, sh in enumerate(shapes)]

    def run(off):
        sims = []
        for q, (sh, eps) in enumerate(zip(shapes, epss)):
            sim = hip().HIPFDTD(sh, DL, npml=4)
            sim.set_epsilon(eps)
            if q in off:
                _index_off(sim)
            assert sim.material_count() == (0 if q in off else 3 + q)
            sims.append(sim)
        from hipfdtd.batch import BatchedFDTD

        b = BatchedFDTD(sims)
        b.set_path("split")
        box = [(s[0] // 2, s[0] // 2 + 1, s[1] // 2, s[1] // 2 + 1, s[2] // 2, s[2] // 2 + 1)
               for s in shapes]  # fmt: skip
        b.set_sources([("z", bx) for bx in box],
                      [lambda t, f=f: fo.gaussian_sine(t, f, 0.5 * f) for f in (2e10, 3e10, 4e10)],
                      24)  # fmt: skip
        b.run(24, graph_steps=8)
        out = [s.fields() for s in sims]
        b.close()
        for s in sims:
            s.close()
        return out

    got, want = run({1}), run({0, 1, 2})
    for q in range(len(shapes)):
        _same(got[q], want[q])
        assert float(np.max(np.abs(got[q]["Ez"]))) > 0


def test_hand_written_coefficients_turn_the_index_off():
    shape = (45, 35, 100)
    eps = _boxes_eps(shape, 6, 4)
    cex = None

    def poke(sim):
        nonlocal cex
        if cex is None:
            cex = sim.read_volume("ce_x") * np.float32(0.9)
        sim.write_volume("ce_x", cex)

    print(f"[mat-poke] shape={shape}", flush=True)
    for rows in (0, 1, 2):
        _same(_run("fused", shape, 6, eps, 2, 12, poke),
              _run("march", shape, 6, eps, 2, 12, poke, rows=rows, materials=0))  # fmt: skip


@pytest.mark.parametrize("rows,planes", [(0, 0), (1, 24), (2, 24)])
def test_indexed_march_full_size(rows, planes):
    """The benchmark's kind of medium on its grid: 40 boxes on vacuum, 256^3,
    with the engine's own tiling, the tiling the benchmark's autotuner picks
    (one row, 24 planes) and two rows."""
    shape = (256, 256, 256)
    eps = _boxes_eps(shape, 41, 2)
    print(f"[mat-full] shape={shape} nvals={len(np.unique(eps))} rows={rows} planes={planes}",
          flush=True)  # fmt: skip
    _same(_run("fused", shape, 12, eps, 1, 6),
          _run("march", shape, 12, eps, 1, 6, rows=rows, planes=planes, materials=41))  # fmt: skip


def test_fp64_engine_has_no_index():
    shape = (45, 35, 100)
    eps = _boxes_eps(shape, 5, 3)
    sim = hip().HIPFDTD(shape, DL, npml=6, path="march", dtype=np.float64)
    sim.set_epsilon(eps)
    assert sim.material_count() == 0
    sim2 = hip().HIPFDTD(shape, DL, npml=6, path="march")
    sim2.set_epsilon(eps, np.zeros(shape, np.float32))  # lossy: ca volumes, no index
    assert sim2.material_count() == 0


@pytest.mark.parametrize("rows", [0, 1, 2])
def test_default_medium_is_indexed(rows):
    """Before any set_epsilon the medium is eps_r = 1 everywhere: one
    material, whose table entry equals the default ce volumes' dt/eps0."""
    shape = (45, 35, 100)
    print(f"[mat-default] rows={rows}", flush=True)
    out = {}
    for path in ("fused", "march"):
        sim = hip().HIPFDTD(shape, DL, npml=6, path=path)
        if path == "march":
            sim.set_march_tiling(rows, 0, 0, -2)
            assert sim.material_count() == 1
        sim.add_source(lambda f: f.add_source_Jx(50, fo.gaussian_sine(f.t, 3e10, 1.5e10), 5, 40,
                                                 4, 30, rim_taper=True))  # fmt: skip
        sim.run(16)
        out[path] = sim.fields()
    _same(out["fused"], out["march"])
    f64 = hip().HIPFDTD(shape, DL, npml=6, path="march", dtype=np.float64)
    assert f64.material_count() == 0
