"""The march sweep's record of a Huygens box's z faces.

The first buffered FP32 monitor of an engine has its two z faces recorded
by the march sweep (the (Ex, Ey) and (Hx, Hy) pairs its co-location reads at
k and k - 1), and its gather reads them from that record instead of from
the fields.  A second monitor on the same box reads the fields.  Both must
hold the same sums bit for bit after every kind of step and state change:
plain march steps, steps with more sources than the sweep applies (the
fix-up kernel changes E after the sweep), field writes, callbacks that read
fields (the step finishes on the split path), path switches, pair steps,
and boxes whose faces sit on a 64-column tile boundary or at the grid edge.
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


def _engine(shape, path="march", npml=6):
    sim = hip().HIPFDTD(shape, 1e-3, npml=npml, path=path)
    rng = np.random.default_rng(9)
    eps = np.ones(shape, np.float32)
    for _ in range(6):
        lo = [int(rng.integers(npml, n - npml - 8)) for n in shape]
        eps[lo[0]:lo[0] + 8, lo[1]:lo[1] + 8, lo[2]:lo[2] + 8] = float(rng.uniform(1.5, 4))
    sim.set_epsilon(eps)
    return sim


def _monitors(sim, lo, hi, freqs=(2.5e10, 3e10, 3.5e10)):
    c = [0.5 * (a + b) * sim.dl for a, b in zip(lo, hi)]
    s = [(b - a) * sim.dl for a, b in zip(lo, hi)]
    a = hip().HIPNear2FarMonitor(sim, c, s, freqs=list(freqs))
    b = hip().HIPNear2FarMonitor(sim, c, s, freqs=list(freqs))
    return a, b


def _same(a, b):
    for q in range(len(a.freqs)):
        for c in FIELDS:
            np.testing.assert_array_equal(a.read_dft(c, q), b.read_dft(c, q), err_msg=f"{c} f{q}")
    pts = [(0.3, 0.2, 1.0), (0.0, 0.0, -1.0), (1.0, 0.5, 0.1)]
    np.testing.assert_array_equal(a.get_farfields(pts, 1), b.get_farfields(pts, 1))


def _sources(sim, n, read_field=None, kface=14):
    """A Jx sheet and n - 1 point currents on the z face k = kface, so that
    a stale record (E changed on the face after the sweep) would show."""
    nx, ny, nz = sim.Nx, sim.Ny, sim.Nz

    def src(f):
        w = fo.gaussian_sine(f.t, 3e10, 1.5e10)
        if read_field is not None and read_field(f.step_num):
            f.read_point("Ex", nx // 2, ny // 2, nz // 2)  # H half runs first, E on the split path
        f.add_source_Jx(nz // 2, w, nx // 3, 2 * nx // 3, ny // 3, 2 * ny // 3, rim_taper=True)
        for q in range(n - 1):
            f.add_source_J("xyz"[q % 3], (0.3 + 0.1 * q) * w,
                           (nx // 2 + q, nx // 2 + q + 1, ny // 2, ny // 2 + 1, kface, kface + 1))  # fmt: skip

    sim.add_source(src)


@pytest.mark.parametrize("case", ["plain", "fixup", "write", "read", "switch", "pair"])
def test_recorded_z_faces_match_field_reads(case):
    shape = (48, 52, 96)
    print(f"[zrec] {case} shape={shape}", flush=True)
    sim = _engine(shape)
    a, b = _monitors(sim, (10, 9, 14), (38, 44, 80))
    assert a.records_zfaces and not b.records_zfaces
    _sources(sim, 6 if case == "fixup" else 2,
             read_field=(lambda n: n % 5 == 2) if case == "read" else None)  # fmt: skip
    for block in range(6):
        if case == "write" and block in (2, 4):
            ex = sim.Ex
            ex[20:30, 20:30, 14] += np.float32(1e-3)
            sim.Ex = ex
        if case == "switch":
            sim.set_path("split" if block % 2 else "march")
        if case == "pair":
            sim.set_path("pair" if block % 2 else "march")
        sim.run(7)
    _same(a, b)
    assert float(np.max(np.abs(a.read_dft("Ex", 1)))) > 0


@pytest.mark.parametrize("lo,hi,records", [
    ((10, 9, 64), (38, 44, 127), True),   # z0 at a 64-column tile edge (k - 1 in the tile before)
    ((0, 0, 1), (47, 51, 94), True),      # box over the whole grid in x and y
    ((10, 9, 0), (38, 44, 80), False),    # z0 = 0 has no k - 1: fields
    ((10, 9, 30), (38, 44, 31), False)])  # z faces one cell apart: fields
def test_record_geometry(lo, hi, records):
    shape = (48, 52, 140)
    print(f"[zrec-geo] box {lo}..{hi}", flush=True)
    sim = _engine(shape)
    a, b = _monitors(sim, lo, hi)
    assert a.records_zfaces == records and not b.records_zfaces
    _sources(sim, 3, kface=max(lo[2], 1))
    sim.run(40)
    _same(a, b)


def test_record_follows_the_monitor_that_holds_it():
    """Turning buffering off releases the record; the next buffered monitor
    takes it, and still matches a monitor on the same box that reads the
    fields."""
    shape = (40, 40, 72)
    sim = _engine(shape)
    a, b = _monitors(sim, (8, 8, 10), (30, 30, 60))
    assert a.records_zfaces
    assert a._lib.hfdtd_n2f_set_buffering(a._handle, 0) == 0
    assert not a.records_zfaces and not b.records_zfaces
    c, d = _monitors(sim, (6, 7, 12), (33, 31, 58))
    assert c.records_zfaces and not d.records_zfaces
    _sources(sim, 2, kface=12)
    sim.run(33)
    _same(c, d)
