"""A Huygens-box gather held for the next march sweep (HIPFDTD.set_overlap).

A K-buffered near-to-far monitor's gather of step n (when it does not fill
the K-step block) is held by the engine, and the march sweep of step n + 1
runs it in blocks after its own: both read state[p] only, the sweep writes
the other slab and the other z-face record (kept once per slab parity).
Every other call on the engine runs a held gather on its own first.  With
this on and off the results must be equal bit for bit (face DFTs, far
fields, probe DFT, fields): over odd and even step counts, contractions
every 2 and 4 steps, two monitors gathering in one step, far-field and field
reads between runs, field writes, switches to the split path, a row-padded
run, a monitor released between runs, per-volume addressing (whose sweep
runs a held gather on its own first) and a plane DFT monitor beside the box
(whose per-step call does the same).
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import fdtd_oracle as fo

pytestmark = pytest.mark.gpu

FIELDS = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")
PTS = [(0.3, 0.2, 1.0), (0.0, 0.0, -1.0), (1.0, 0.5, 0.1)]


def hip():
    import hipfdtd
    import hipfdtd.dft  # noqa: F401
    import hipfdtd.probes  # noqa: F401

    return hipfdtd


def _run(case, overlap):
    shape = (44, 46, 70) if case == "padded" else (44, 46, 64)
    sim = hip().HIPFDTD(shape, 1e-3, npml=6, path="march")
    rng = np.random.default_rng(5)
    eps = np.ones(shape, np.float32)
    for _ in range(5):
        lo = [int(rng.integers(6, n - 14)) for n in shape]
        eps[lo[0]:lo[0] + 8, lo[1]:lo[1] + 8, lo[2]:lo[2] + 8] = float(rng.uniform(1.5, 4))
    sim.set_epsilon(eps)
    sim.set_row_padding(case == "padded")
    if case == "pervolume":  # the sweep for slabs of 4 GB and more: a held gather runs first
        sim.set_march_addressing(True)
    sim.set_overlap(overlap)
    lo, hi = (9, 8, 12), (34, 37, 52)
    c = [0.5 * (a + b) * sim.dl for a, b in zip(lo, hi)]
    s = [(b - a) * sim.dl for a, b in zip(lo, hi)]
    freqs = [2.5e10, 3e10, 3.5e10]
    a = hip().HIPNear2FarMonitor(sim, c, s, freqs=freqs, buffer_steps=4)
    b = hip().HIPNear2FarMonitor(sim, c, s, freqs=freqs)
    assert a.records_zfaces and not b.records_zfaces
    pr = hip().probes.HIPPointProbe(sim, (20, 21, 30), "Ez", freqs)
    extra = hip().HIPNear2FarMonitor(sim, c, s, freqs=freqs[:1], buffer_steps=2)
    # a plane DFT's per-step call runs a held gather before its own
    pl = (hip().dft.HIPPlaneDFTMonitor(sim, "z", 30, ("Ex", "Hy"), freqs[:2])
          if case == "plane" else None)  # fmt: skip
    nx, ny, nz = shape
    sim.add_source(lambda f: f.add_source_Jx(nz // 2, fo.gaussian_sine(f.t, 3e10, 1.5e10),
                                             nx // 3, 2 * nx // 3, ny // 3, 2 * ny // 3))  # fmt: skip
    if case == "padded":
        assert sim._padded_begin()
        hip()._lib.check(sim._lib.hfdtd_padded_end(sim._handle))
    out = {}
    for block in range(6):
        if case == "switch":
            sim.set_path("split" if block % 3 == 1 else "march")
        if case == "write" and block in (2, 4):
            ex = sim.Ex
            ex[20:26, 20:26, 12] += np.float32(1e-3)
            sim.Ex = ex
        if case == "drop" and block == 3:  # released with gathers still queued
            sim.clear_monitors()
            extra._release()
            for m in (a, b, pr):
                sim.add_monitor(m)
        sim.run(7 + block)
        if case == "midread" and block in (1, 4):
            out[f"far{block}"] = a.get_farfields(PTS, 1)
            out[f"ez{block}"] = sim.read_point("Ez", nx // 2, ny // 2, nz // 2)
    for k in range(len(freqs)):
        for n in FIELDS:
            out[f"a{n}{k}"] = a.read_dft(n, k)
            out[f"b{n}{k}"] = b.read_dft(n, k)
    out["fa"] = a.get_farfields(PTS, 0)
    out["fb"] = b.get_farfields(PTS, 2)
    out["probe"] = pr.dft()
    if pl is not None:
        out["plane"] = pl.dft()
    if case != "drop":
        out["extra"] = extra.read_dft("Hy", 0)
    out.update({"f" + n: v for n, v in sim.fields().items()})
    sim.close()
    return out


@pytest.mark.parametrize("case", ["plain", "midread", "write", "switch", "padded", "drop",
                                  "pervolume", "plane"])
def test_overlapped_gather_matches_serial(case):
    print(f"[overlap] {case}", flush=True)
    want = _run(case, False)
    got = _run(case, True)
    assert set(got) == set(want)
    for k in want:
        np.testing.assert_array_equal(got[k], want[k], err_msg=k)
    assert float(np.max(np.abs(want["aEx1"]))) > 0


def test_overlap_knob_rejects_bad_values():
    sim = hip().HIPFDTD((16, 16, 16), 1e-3, npml=2, path="march")
    assert sim._lib.hfdtd_set_overlap(sim._handle, 2) != 0
    assert sim._lib.hfdtd_set_overlap(None, 1) != 0
    sim.close()
