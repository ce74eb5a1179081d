"""GPU: batched sweep of uneven grids and the device point probe.

The batched sweep must reproduce, grid by grid, what the plain engine gives
when each simulation is stepped on its own (same sources, same probes):
fields within 1e-6 of the field-group peak and probe DFTs within 1e-6
relative; the serial engine itself is pinned to the oracle elsewhere.
"""

from __future__ import annotations

import math

import numpy as np
import pytest

from oracle import fdtd_oracle as fo

pytestmark = pytest.mark.gpu

FIELDS = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")
C0 = 299_792_458.0


def hip():
    import hipfdtd
    import hipfdtd.batch
    import hipfdtd.probes

    return hipfdtd


def group_error(a, b):
    worst = 0.0
    for grp in (FIELDS[:3], FIELDS[3:]):
        peak = max(float(np.max(np.abs(b[n]))) for n in grp)
        for n in grp:
            worst = max(worst, float(np.max(np.abs(a[n].astype(np.float64) - b[n]))) / peak)
    return worst


def sphere_case(rng, shape):
    eps = np.ones(shape, np.float32)
    r = int(rng.integers(3, max(4, min(shape) // 4)))
    c = [s // 2 for s in shape]
    I, J, K = np.indices(shape)
    eps[(I - c[0]) ** 2 + (J - c[1]) ** 2 + (K - c[2]) ** 2 < r * r] = float(rng.uniform(1.5, 8))
    f = float(rng.uniform(C0 / (30 * 1e-3), C0 / (12 * 1e-3)))
    return eps, f


def waveform(f):
    fw = 0.5 * f
    return lambda t: fo.gaussian_sine(t, f, fw)


def test_point_probe_matches_host_dft():
    shape = (24, 26, 28)
    sim = hip().HIPFDTD(shape, 1e-3, npml=4)
    ora = fo.OracleFDTD(shape, 1e-3, npml=4)
    wf = waveform(2e10)
    freqs = [1.5e10, 2e10, 2.5e10]
    probe = hip().probes.HIPPointProbe(sim, (14, 9, 17), "Ez", freqs)
    acc = np.zeros(3, complex)
    om = 2 * np.pi * np.array(freqs)
    for s in (sim, ora):
        s.add_source(lambda f: f.add_source_J("z", wf(f.t), (12, 13, 13, 14, 14, 15)))
    for _ in range(150):
        sim.step()
        ora.step()
        acc += float(ora.Ez[14, 9, 17]) * np.exp(1j * om * ora.t) * ora.dt
    got = probe.dft()
    assert np.max(np.abs(got - acc)) / np.max(np.abs(acc)) < 1e-5


@pytest.mark.parametrize("graph_steps,path", [(0, "fused"), (10, "fused"), (0, "split"),
                                              (10, "split")])  # fmt: skip
def test_batch_matches_serial_engines(graph_steps, path):
    """Batched fused tiles match the serial fused engines to 1e-6 (the sweep
    differs from the serial engines' launch shape only); the batched split
    sweeps match serial split engines bit for bit."""
    rng = np.random.default_rng(4)
    shapes = [tuple(int(v) for v in rng.integers(20, 72, size=3)) for _ in range(6)]
    print(f"[batch] shapes={shapes} graph_steps={graph_steps} path={path}", flush=True)
    n_steps = 60
    cases = [sphere_case(rng, s) for s in shapes]
    nf = 4

    def src_box(shape):
        c = [s // 2 for s in shape]
        return (c[0], c[0] + 1, c[1], c[1] + 1, c[2], c[2] + 1)

    def probe_cell(shape):
        return (shape[0] // 2 + 3, shape[1] // 2 - 2, shape[2] // 2 + 1)

    freqs = [list(np.linspace(0.7, 1.3, nf) * f) for _, f in cases]
    serial = []
    for shape, (eps, f), fr in zip(shapes, cases, freqs):
        sim = hip().HIPFDTD(shape, 1e-3, npml=8, fused=path == "fused")
        sim.set_epsilon(eps)
        wf = waveform(f)
        sim.add_source(lambda s, wf=wf, b=src_box(shape): s.add_source_J("z", wf(s.t), b))
        pr = hip().probes.HIPPointProbe(sim, probe_cell(shape), "Ez", fr)
        sim.run(n_steps)
        serial.append((sim.fields(), pr.dft()))
    sims = []
    for shape, (eps, f) in zip(shapes, cases):
        sim = hip().HIPFDTD(shape, 1e-3, npml=8)
        sim.set_epsilon(eps)
        sims.append(sim)
    batch = hip().batch.BatchedFDTD(sims)
    batch.set_path(path)
    batch.set_sources([("z", src_box(s)) for s in shapes], [waveform(f) for _, f in cases],
                      n_steps)  # fmt: skip
    batch.set_probes([probe_cell(s) for s in shapes], "Ez", freqs, n_steps)
    batch.run(n_steps // 2, graph_steps)
    batch.run(n_steps - n_steps // 2, graph_steps)
    dfts = batch.probe_dfts()
    for q, sim in enumerate(sims):
        ref_f, ref_p = serial[q]
        if path == "split":
            for name in ref_f:
                np.testing.assert_array_equal(sim.fields()[name], ref_f[name], err_msg=name)
            np.testing.assert_array_equal(dfts[q], ref_p)
        assert group_error(sim.fields(), ref_f) < 1e-6, q
        assert np.max(np.abs(dfts[q] - ref_p)) / np.max(np.abs(ref_p)) < 1e-6, q
        assert sim.step_num == n_steps and math.isclose(sim.t, n_steps * sim.dt, rel_tol=1e-12)
    batch.close()


def _march_batch_case(shapes, media, n_steps, graph_steps, rows, planes, run, fp64=False,
                      lossy=(), split=None, width=0, pad=None):
    """Serial march engines and one batched march run of the same grids,
    sources and probes; returns [(serial fields, serial probe DFT)], the
    batch's sims and its probe DFTs."""
    dtype = np.float64 if fp64 else np.float32
    nf = 4
    freqs = [list(np.linspace(0.7, 1.3, nf) * f) for _, f in media]

    def src_box(shape):
        c = [s // 2 for s in shape]
        return (c[0], c[0] + 1, c[1], c[1] + 1, c[2], c[2] + 1)

    def probe_cell(shape):
        return (shape[0] // 2 + 3, shape[1] // 2 - 2, shape[2] // 2 + 1)

    def engine(q, shape, eps):
        sim = hip().HIPFDTD(shape, 1e-3, npml=8, dtype=dtype, path="march")
        if q in lossy:
            sig = np.zeros(shape, np.float32)
            sig[2:-2, 3:-3, 4:-4] = 0.05
            sim.set_epsilon(eps, sig)
        else:
            sim.set_epsilon(eps)
        return sim

    serial = []
    for q, (shape, (eps, f), fr) in enumerate(zip(shapes, media, freqs)):
        sim = engine(q, shape, eps)
        wf = waveform(f)
        sim.add_source(lambda s, wf=wf, b=src_box(shape): s.add_source_J("z", wf(s.t), b))
        pr = hip().probes.HIPPointProbe(sim, probe_cell(shape), "Ez", fr)
        sim.run(n_steps)
        serial.append((sim.fields(), pr.dft()))
        sim.close()
    sims = [engine(q, shape, eps) for q, (shape, (eps, _)) in enumerate(zip(shapes, media))]
    batch = hip().batch.BatchedFDTD(sims)
    batch.set_path("march")
    batch.set_march(rows, planes, run, width)
    if pad is not None:
        batch.set_march_padding(pad)
    batch.set_sources([("z", src_box(s)) for s in shapes], [waveform(f) for _, f in media],
                      n_steps)  # fmt: skip
    batch.set_probes([probe_cell(s) for s in shapes], "Ez", freqs, n_steps)
    first = n_steps // 2 + 1 if split is None else split  # an odd first run: parity carries over
    batch.run(first, graph_steps)
    batch.run(n_steps - first, graph_steps)
    dfts = batch.probe_dfts()
    batch.close()
    return serial, sims, dfts


@pytest.mark.parametrize("pad,width,graph_steps", [(True, 64, 10), (False, 64, 10),
                                                    (True, 32, 0), (True, 64, 7)])  # fmt: skip
def test_batch_march_row_padding(pad, width, graph_steps):
    """Engines whose nz is not a multiple of 32 run on copies with padded
    rows (material-indexed FP32 batches): fields and probe DFTs equal the
    serial march engines bit for bit with the padding on and off, over two
    runs whose first has an odd step count (the copies are written back at
    the parity the run ends on)."""
    rng = np.random.default_rng(17)
    shapes = [(40, 37, 33), (29, 44, 47), (36, 30, 64), (31, 35, 95), (27, 41, 100)]
    media = [sphere_case(rng, sh) for sh in shapes]
    print(f"[batch-pad] pad={pad} width={width} graph_steps={graph_steps}", flush=True)
    serial, sims, dfts = _march_batch_case(shapes, media, 31, graph_steps, 1, 16, 16,
                                           width=width, pad=pad)  # fmt: skip
    for q, sim in enumerate(sims):
        assert sim.material_count() > 0
        ref_f, ref_p = serial[q]
        got = sim.fields()
        for n in FIELDS:
            np.testing.assert_array_equal(got[n], ref_f[n], err_msg=f"grid {q} {n}")
        np.testing.assert_array_equal(dfts[q], ref_p, err_msg=f"grid {q} probe")
        assert sim.step_num == 31
        sim.run(2)  # the engine steps on from the written-back state
        sim.close()


@pytest.mark.parametrize("medium,graph_steps,rows,planes,run", [
    ("index", 0, 1, 32, 16), ("index", 10, 2, 24, 8), ("index", 10, 1, 8, 1),
    ("volumes", 10, 1, 30, 16), ("lossy", 0, 2, 32, 16), ("lossy", 10, 1, 64, 5),
    ("fp64", 10, 1, 32, 16), ("fp64-lossy", 0, 1, 16, 16)])  # fmt: skip
def test_batch_march_matches_serial_march_engines(medium, graph_steps, rows, planes, run):
    """The batched march sweep against the same grids stepped one by one on
    the single-engine march sweep: fields and probe DFTs bit for bit, on
    the material index (every grid indexed), the ce volumes (one grid with
    more than 48 eps values), with ca (one lossy grid) and in FP64."""
    rng = np.random.default_rng(7)
    shapes = [tuple(int(v) for v in rng.integers(20, 90, size=3)) for _ in range(5)]
    media = [sphere_case(rng, s) for s in shapes]
    if medium == "volumes":
        eps = media[1][0].copy()
        eps[2:-2, 2:-2, 2:-2] += rng.uniform(0, 1e-3, size=eps[2:-2, 2:-2, 2:-2].shape)
        media[1] = (eps.astype(np.float32), media[1][1])
    lossy = (2,) if "lossy" in medium else ()
    n_steps = 41
    print(f"[batch-march] {medium} shapes={shapes} graph={graph_steps} rows={rows} "
          f"planes={planes} run={run}", flush=True)  # fmt: skip
    serial, sims, dfts = _march_batch_case(shapes, media, n_steps, graph_steps, rows, planes, run,
                                           fp64=medium.startswith("fp64"), lossy=lossy)  # fmt: skip
    for q, sim in enumerate(sims):
        if medium == "index":
            assert sim.material_count() > 0
        ref_f, ref_p = serial[q]
        got = sim.fields()
        for name in ref_f:
            np.testing.assert_array_equal(got[name], ref_f[name], err_msg=f"grid {q} {name}")
        np.testing.assert_array_equal(dfts[q], ref_p, err_msg=f"grid {q} probe")
        assert np.max(np.abs(ref_p)) > 0
        assert sim.step_num == n_steps and math.isclose(sim.t, n_steps * sim.dt, rel_tol=1e-12)
        sim.close()


@pytest.mark.parametrize("fp64,lossy,graph_steps,runs", [
    (True, False, 0, 1), (True, True, 10, 4), (False, True, 10, 3), (False, False, 10, 8)])  # fmt: skip
def test_batch_split_fp64_and_lossy_match_serial(fp64, lossy, graph_steps, runs):
    """The split batch (sources and probes in the hook pass over their
    planes) on FP64 engines, with a lossy grid (the kernels with ca) and with
    1 to 8 runs of 256 cells per work item: fields and probe DFTs bit for bit
    against serial split engines."""
    rng = np.random.default_rng(21)
    shapes = [tuple(int(v) for v in rng.integers(18, 60, size=3)) for _ in range(4)]
    media = [sphere_case(rng, s) for s in shapes]
    dtype = np.float64 if fp64 else np.float32
    n_steps, nf = 37, 3
    freqs = [list(np.linspace(0.8, 1.2, nf) * f) for _, f in media]

    def box(shape):
        c = [v // 2 for v in shape]
        return (c[0] - 1, c[0] + 1, c[1], c[1] + 2, c[2], c[2] + 1)

    def cell(shape):
        return (shape[0] // 2 - 2, shape[1] // 2 + 1, shape[2] // 2 + 2)

    def engine(q, shape, eps):
        sim = hip().HIPFDTD(shape, 1e-3, npml=6, dtype=dtype, path="split")
        if lossy and q == 1:
            sig = np.zeros(shape, np.float32)
            sig[3:-3, 2:-2, 4:-4] = 0.08
            sim.set_epsilon(eps, sig)
        else:
            sim.set_epsilon(eps)
        return sim

    serial = []
    for q, (shape, (eps, f), fr) in enumerate(zip(shapes, media, freqs)):
        sim = engine(q, shape, eps)
        wf = waveform(f)
        sim.add_source(lambda s, wf=wf, b=box(shape): s.add_source_J("y", wf(s.t), b))
        pr = hip().probes.HIPPointProbe(sim, cell(shape), "Ey", fr)
        sim.run(n_steps)
        serial.append((sim.fields(), pr.dft()))
        sim.close()
    sims = [engine(q, shape, eps) for q, (shape, (eps, _)) in enumerate(zip(shapes, media))]
    batch = hip().batch.BatchedFDTD(sims)
    batch.set_path("split")
    batch.set_split(runs)
    batch.set_sources([("y", box(s)) for s in shapes], [waveform(f) for _, f in media], n_steps)
    batch.set_probes([cell(s) for s in shapes], "Ey", freqs, n_steps)
    batch.run(15, graph_steps)
    batch.run(n_steps - 15, graph_steps)
    dfts = batch.probe_dfts()
    batch.close()
    for q, sim in enumerate(sims):
        ref_f, ref_p = serial[q]
        got = sim.fields()
        for name in ref_f:
            np.testing.assert_array_equal(got[name], ref_f[name], err_msg=f"grid {q} {name}")
        np.testing.assert_array_equal(dfts[q], ref_p, err_msg=f"grid {q} probe")
        assert np.max(np.abs(ref_p)) > 0
        sim.close()


def test_batch_march_against_oracle():
    """Two grids batched on the march path against the oracle stepping each
    grid on its own (fields and probe DFT within 1e-5)."""
    rng = np.random.default_rng(9)
    shapes = [(30, 41, 70), (52, 27, 33)]
    media = [sphere_case(rng, s) for s in shapes]
    n_steps = 60
    _, sims, dfts = _march_batch_case(shapes, media, n_steps, 0, 1, 16, 16)
    for q, (shape, (eps, f)) in enumerate(zip(shapes, media)):
        ora = fo.OracleFDTD(shape, 1e-3, npml=8)
        ora.set_epsilon(eps)
        c = [s // 2 for s in shape]
        box = (c[0], c[0] + 1, c[1], c[1] + 1, c[2], c[2] + 1)
        wf = waveform(f)
        ora.add_source(lambda s, wf=wf, b=box: s.add_source_J("z", wf(s.t), b))
        cell = (shape[0] // 2 + 3, shape[1] // 2 - 2, shape[2] // 2 + 1)
        om = 2 * np.pi * np.linspace(0.7, 1.3, 4) * f
        acc = np.zeros(4, complex)
        for _ in range(n_steps):
            ora.step()
            acc += float(ora.Ez[cell]) * np.exp(1j * om * ora.t) * ora.dt
        assert group_error(sims[q].fields(), ora.fields()) < 1e-5, q
        assert np.max(np.abs(dfts[q] - acc)) / np.max(np.abs(acc)) < 1e-5, q
        sims[q].close()


def test_batch_march_rejects_bad_shapes():
    sims = [hip().HIPFDTD((16, 16, 16), 1e-3, npml=2) for _ in range(2)]
    batch = hip().batch.BatchedFDTD(sims)
    with pytest.raises(ValueError, match="rows"):
        batch.set_march(rows=3)
    with pytest.raises(ValueError, match="width"):
        batch.set_march(width=16)
    with pytest.raises(ValueError, match="runs"):
        batch.set_split(9)
    with pytest.raises(ValueError, match="64-column"):
        batch.set_march(rows=2, width=32)
    with pytest.raises(ValueError, match="path"):
        batch.set_path("pair")
    batch.close()
    s64 = [hip().HIPFDTD((16, 16, 16), 1e-3, npml=2, dtype=np.float64) for _ in range(2)]
    b64 = hip().batch.BatchedFDTD(s64)
    with pytest.raises(ValueError, match="FP64"):
        b64.set_march(rows=2)
    b64.close()


def test_batch_rejects_mixed_grids():
    a = hip().HIPFDTD((16, 16, 16), 1e-3, npml=2)
    b = hip().HIPFDTD((16, 16, 16), 2e-3, npml=2)
    with pytest.raises(ValueError, match="same dl"):
        hip().batch.BatchedFDTD([a, b])


@pytest.mark.parametrize("shape,path,graph_steps,dtype", [
    ((40, 36, 44), "split", 16, np.float32), ((48, 40, 70), "split", 0, np.float32),
    ((64, 64, 64), "march", 10, np.float32), ((40, 36, 44), "fused", 16, np.float32),
    ((36, 30, 50), "fused", 10, np.float64), ((44, 40, 48), "pair", 8, np.float32)])  # fmt: skip
def test_run_device_matches_eager_run(shape, path, graph_steps, dtype):
    """HIPFDTD.run_device (a graph-replayed batch of one, source and probe on
    the device) against run() with the same source as a callback and a
    point-probe monitor: fields and the probe DFT bit for bit; the engine's
    time and step count advance alike, and stepping continues normally."""
    rng = np.random.default_rng(31)
    eps, f = sphere_case(rng, shape)
    c = [v // 2 for v in shape]
    box = (c[0], c[0] + 1, c[1] - 1, c[1] + 1, c[2], c[2] + 1)
    cell = (c[0] + 2, c[1] - 3, c[2] + 1)
    freqs = [0.8 * f, f, 1.2 * f]
    wf = waveform(f)
    n_steps = 33
    print(f"[run-device] shape={shape} path={path} graph_steps={graph_steps} "
          f"dtype={np.dtype(dtype).name}", flush=True)  # fmt: skip
    ref = hip().HIPFDTD(shape, 1e-3, npml=6, path=path, dtype=dtype)
    ref.set_epsilon(eps)
    ref.add_source(lambda s: s.add_source_J("x", wf(s.t), box))
    pr = hip().probes.HIPPointProbe(ref, cell, "Ex", freqs)
    ref.run(n_steps)
    want_f, want_p = ref.fields(), pr.dft()
    sim = hip().HIPFDTD(shape, 1e-3, npml=6, path=path, dtype=dtype)
    sim.set_epsilon(eps)
    got_p = sim.run_device(n_steps, ("x", box, wf), (cell, "Ex", freqs), graph_steps=graph_steps)
    got_f = sim.fields()
    for n in FIELDS:
        np.testing.assert_array_equal(got_f[n], want_f[n], err_msg=n)
    np.testing.assert_array_equal(got_p, want_p)
    assert sim.step_num == ref.step_num and sim.t == ref.t
    sim.run(3)  # the engine steps on after a device run
    assert sim.step_num == n_steps + 3
    with pytest.raises(ValueError, match="callbacks"):
        ref.run_device(2, ("x", box, wf))
    ref.close()
    sim.close()
