"""Row-padded march runs (hfdtd_padded_begin / end, used by HIPFDTD.run).

Where nz does not fill whole 128-byte lines, run() on the march path without
monitors steps a copy whose rows are padded to whole lines.  The padding only
moves where each row starts, so fields must equal those of the same run on
the dense layout bit for bit: on every medium kind, in FP64, on per-volume
addressing, with more sources than the sweep applies (the run ends and the
step takes the fix-up kernel), with a source callback that reads the fields
mid-run (the run ends before the read), and across several run() calls.
"""

from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from oracle import fdtd_oracle as fo

pytestmark = pytest.mark.gpu

FIELDS = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")


def hip():
    import hipfdtd

    return hipfdtd


def _medium(kind, shape, seed=3):
    rng = np.random.default_rng(seed)
    sig = None
    if kind == "index":
        eps = np.ones(shape, np.float32)
        eps[3:-4, 5:-3, 6:-5] = 2.5
        eps[shape[0] // 2:, :, : shape[2] // 3] = 4.0
    else:
        eps = rng.uniform(1.0, 5.0, size=shape).astype(np.float32)
        if kind == "lossy":
            sig = rng.uniform(0.0, 1.5, size=shape).astype(np.float32)
    return eps, sig


def _run(shape, kind, pad, *, dtype=np.float32, nsrc=2, steps=(9, 8), per_volume=False,
         reader=False):  # fmt: skip
    sim = hip().HIPFDTD(shape, 1e-3, npml=5, dtype=dtype, path="march")
    eps, sig = _medium(kind, shape)
    sim.set_epsilon(eps, sig)
    if per_volume:
        sim.set_march_addressing(True)
    sim.set_row_padding(pad)
    nx, ny, nz = shape
    seen = []

    def src(f):
        a = fo.gaussian_sine(f.t, 3e10, 1.5e10)
        f.add_source_Jx(nz // 2, a, 2, nx - 3, 3, ny - 2, rim_taper=True, rim_edge=0.7)
        for q in range(nsrc - 1):
            f.add_source_J("xyz"[q % 3], (0.3 - 0.05 * q) * a,
                           (nx // 3 + q, nx // 3 + q + 1, ny // 2 - q, ny // 2 - q + 1,
                            5 + 2 * q, 6 + 2 * q))  # fmt: skip
        if reader and f.step_num == 5:
            seen.append(float(f.read_point("Ez", nx // 2, ny // 2, nz // 2)))

    sim.add_source(src)
    if pad:  # the run takes the padded copy (eligible: march path, no monitors, nz off-line)
        assert sim._padded_begin()
        hip()._lib.check(sim._lib.hfdtd_padded_end(sim._handle))
    for n in steps:
        sim.run(n)
    out = sim.fields()
    live = C.c_int(-1)
    hip()._lib.check(sim._lib.hfdtd_padded_live(sim._handle, C.byref(live)))
    assert live.value == 0  # run() always ends the padded run
    sim.close()
    return out, seen


@pytest.mark.parametrize("shape,kind,dtype,nsrc,per_volume,reader", [
    ((36, 30, 45), "index", np.float32, 2, False, False),
    ((33, 41, 70), "eps", np.float32, 3, False, False),
    ((40, 28, 50), "lossy", np.float32, 2, False, False),
    ((30, 34, 37), "eps", np.float64, 2, False, False),
    ((31, 29, 66), "index", np.float32, 6, False, False),
    ((34, 36, 47), "index", np.float32, 2, True, False),
    ((29, 33, 55), "eps", np.float32, 2, False, True)])  # fmt: skip
def test_padded_run_matches_dense_run(shape, kind, dtype, nsrc, per_volume, reader):
    print(f"[padded] shape={shape} {kind} {np.dtype(dtype).name} nsrc={nsrc} "
          f"per_volume={per_volume} reader={reader}", flush=True)  # fmt: skip
    want, seen_w = _run(shape, kind, False, dtype=dtype, nsrc=nsrc, per_volume=per_volume,
                        reader=reader)  # fmt: skip
    got, seen_g = _run(shape, kind, True, dtype=dtype, nsrc=nsrc, per_volume=per_volume,
                       reader=reader)  # fmt: skip
    for n in FIELDS:
        np.testing.assert_array_equal(got[n], want[n], err_msg=n)
    assert seen_g == seen_w


def test_padded_run_through_the_c_abi():
    """begin / live / end around plain march steps; any other call ends the
    run; a grid whose rows fill whole lines takes no copy."""
    L = hip()._lib
    shape = (26, 30, 41)
    a = hip().HIPFDTD(shape, 1e-3, npml=4, path="march")
    b = hip().HIPFDTD(shape, 1e-3, npml=4, path="march")
    src = (L.Source * 1)(L.Source(2, L.SRC_CURRENT, 13, 14, 15, 16, 20, 21, 0, 0.0, 0.7))
    live = C.c_int(0)
    L.check(a._lib.hfdtd_padded_begin(a._handle))
    L.check(a._lib.hfdtd_padded_live(a._handle, C.byref(live)))
    assert live.value == 1
    for _ in range(7):
        L.check(a._lib.hfdtd_step(a._handle, src, 1, L.PATH_MARCH))
        L.check(b._lib.hfdtd_step(b._handle, src, 1, L.PATH_MARCH))
    ez_a = a.Ez  # a read ends the run
    L.check(a._lib.hfdtd_padded_live(a._handle, C.byref(live)))
    assert live.value == 0
    np.testing.assert_array_equal(ez_a, b.Ez)
    L.check(a._lib.hfdtd_padded_end(a._handle))  # ending twice is a no-op
    a.close()
    b.close()
    c = hip().HIPFDTD((20, 22, 64), 1e-3, npml=3, path="march")
    L.check(c._lib.hfdtd_padded_begin(c._handle))
    L.check(c._lib.hfdtd_padded_live(c._handle, C.byref(live)))
    assert live.value == 0  # 64 floats: rows already fill whole lines
    c.close()


@pytest.mark.parametrize("buffer_steps,accum,probe,plane", [
    (None, None, True, False), (0, "fp64", False, False), (0, "kahan", True, False),
    (None, None, False, True)])  # fmt: skip
def test_padded_run_with_box_monitor_and_probe(buffer_steps, accum, probe, plane):
    """A Huygens box (its z faces recorded by the sweep when buffered) and a
    point probe sample the padded copy: face DFTs, far field and probe DFT
    equal the dense run's bit for bit.  With a plane DFT monitor the run keeps
    the dense layout, with the same results."""
    import hipfdtd.dft
    import hipfdtd.probes

    shape = (38, 42, 53)
    print(f"[padded-mon] K={buffer_steps} accum={accum} probe={probe} plane={plane}", flush=True)
    res = []
    for pad in (False, True):
        sim = hip().HIPFDTD(shape, 1e-3, npml=5, path="march")
        eps, _ = _medium("index", shape)
        sim.set_epsilon(eps)
        sim.set_row_padding(pad)
        lo, hi = (9, 10, 12), (29, 31, 40)
        c = [0.5 * (a + b) * sim.dl for a, b in zip(lo, hi)]
        size = [(b - a) * sim.dl for a, b in zip(lo, hi)]
        kw = {} if buffer_steps is None else {"buffer_steps": buffer_steps}
        mon = hip().HIPNear2FarMonitor(sim, c, size, freqs=[2.5e10, 3e10], accum=accum, **kw)
        pr = hip().probes.HIPPointProbe(sim, (20, 21, 30), "Ez", [2.5e10, 3e10]) if probe else None
        pl = (hip().dft.HIPPlaneDFTMonitor(sim, "z", 26, ("Ex", "Ey"), [3e10])
              if plane else None)  # fmt: skip
        if pad:
            assert sim._padded_begin() == (not plane)
            hip()._lib.check(sim._lib.hfdtd_padded_end(sim._handle))
        nx, ny, nz = shape
        sim.add_source(lambda f: f.add_source_Jx(nz // 2, fo.gaussian_sine(f.t, 3e10, 1.5e10),
                                                 4, nx - 4, 4, ny - 4))  # fmt: skip
        sim.run(23)
        sim.run(18)
        out = {k: mon.read_dft(k, 1) for k in FIELDS}
        out["far"] = mon.get_farfields([(0.3, 0.2, 1.0), (0.0, 0.0, -1.0)], 0)
        if pr is not None:
            out["probe"] = pr.dft()
        if pl is not None:
            out["plane"] = pl.dft()
        out.update({"f" + k: v for k, v in sim.fields().items()})
        res.append(out)
        sim.close()
    for k in res[0]:
        np.testing.assert_array_equal(res[1][k], res[0][k], err_msg=k)
