"""GPU: multi-frequency plane DFT (per-step and K-buffered matrix-core forms).

Every mode is compared with an FP64 DFT of the same device field values read
back each step (host side, same phase recurrence), and the buffered result
with a DFT of the oracle's fields on the same inputs.  Tolerances are
relative to the largest |sum| of the plane: FP64 sums 1e-12 (same samples,
same order up to FP64 rounding), compensated FP32 2e-7, plain FP32 and the
buffered FP32 contraction 1e-5; against the oracle 1e-5 (field parity).
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import fdtd_oracle as fo

pytestmark = pytest.mark.gpu

SHAPE = (24, 26, 28)
NPML = 4
DL = 1e-3
F0 = 2.0e10


def hip():
    import hipfdtd
    import hipfdtd.dft

    return hipfdtd


def waveform(t):
    return fo.gaussian_sine(t, F0, 0.5 * F0)


def drive(sim):
    sim.add_source(lambda f: f.add_source_J("z", waveform(f.t), (11, 12, 13, 14, 13, 14)))
    sim.add_source(lambda f: f.add_source_J("x", 0.5 * waveform(f.t), (15, 16, 9, 10, 17, 18)))


def rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


class HostPlaneDFT:
    """FP64 reference: reads the plane of each component every step."""

    def __init__(self, sim, axis, index, comps, freqs):
        self.axis, self.index, self.comps = axis, index, comps
        self.om = 2 * np.pi * np.asarray(freqs)
        self.acc = None
        self.ph = None
        self.d = np.exp(1j * self.om * sim.dt)
        sim.add_monitor(self)

    def plane(self, vol):
        return {"x": vol[self.index], "y": vol[:, self.index], "z": vol[:, :, self.index]}[self.axis]

    def __call__(self, f):
        self.ph = np.exp(1j * self.om * f.t) * f.dt if self.ph is None else self.ph * self.d
        vals = np.stack([self.plane(getattr(f, c)).astype(np.float64) for c in self.comps])
        term = vals[..., None] * self.ph
        self.acc = term if self.acc is None else self.acc + term


@pytest.mark.parametrize("steps", [130, 64])
def test_plane_dft_modes_match_fp64_host_sums(steps):
    print(f"[dft-modes] steps={steps}", flush=True)
    freqs = list(np.linspace(0.6, 1.4, 16) * F0)
    sim = hip().HIPFDTD(SHAPE, DL, npml=NPML)
    drive(sim)
    P = hip().dft.HIPPlaneDFTMonitor
    mons = {
        "buf8": P(sim, "z", 14, ("Ex", "Ey"), freqs, mode="buffered", K=8),
        "buf32": P(sim, "z", 14, ("Ex", "Ey"), freqs, mode="buffered", K=32),
        "buf32_kahan": P(sim, "z", 14, ("Ex", "Ey"), freqs, mode="buffered", K=32, accum="kahan"),
        "buf32_fp64": P(sim, "z", 14, ("Ex", "Ey"), freqs, mode="buffered", K=32, accum="fp64"),
        "fp32": P(sim, "z", 14, ("Ex", "Ey"), freqs, mode="direct", accum="fp32"),
        "kahan": P(sim, "z", 14, ("Ex", "Ey"), freqs, mode="direct", accum="kahan"),
        "fp64": P(sim, "z", 14, ("Ex", "Ey"), freqs, mode="direct", accum="fp64"),
    }
    xmon = P(sim, "x", 9, ("Ez", "Hx", "Hz"), freqs[:5], mode="buffered", K=16)
    ymon = P(sim, "y", 17, ("Ey",), freqs[:3], mode="buffered", K=6)
    host = HostPlaneDFT(sim, "z", 14, ("Ex", "Ey"), freqs)
    hx = HostPlaneDFT(sim, "x", 9, ("Ez", "Hx", "Hz"), freqs[:5])
    hy = HostPlaneDFT(sim, "y", 17, ("Ey",), freqs[:3])
    sim.run(steps)
    ref = host.acc
    errs = {k: rel(m.dft(), ref) for k, m in mons.items()}
    print("[dft-modes] errors vs host FP64:", errs, flush=True)
    assert errs["fp64"] < 1e-12
    assert errs["kahan"] < 2e-7
    assert errs["fp32"] < 1e-5
    for k in ("buf8", "buf32", "buf32_kahan", "buf32_fp64"):
        assert errs[k] < 1e-5, k
    assert rel(xmon.dft(), hx.acc) < 1e-5
    assert rel(ymon.dft(), hy.acc) < 1e-5
    assert mons["buf8"].dft().shape == (2, SHAPE[0], SHAPE[1], 16)
    assert xmon.dft().shape == (3, SHAPE[1], SHAPE[2], 5)


def test_plane_dft_buffered_matches_oracle_fields():
    freqs = [0.8 * F0, F0, 1.2 * F0]
    eps = np.ones(SHAPE, np.float32)
    eps[6:18, 8:20, 5:15] = 3.0
    sim = hip().HIPFDTD(SHAPE, DL, npml=NPML)
    sim.set_epsilon(eps)
    drive(sim)
    mon = hip().dft.HIPPlaneDFTMonitor(sim, "z", 20, ("Ex", "Ez"), freqs, K=32)
    ora = fo.OracleFDTD(SHAPE, DL, npml=NPML)
    ora.set_epsilon(eps)
    drive(ora)
    host = HostPlaneDFT(ora, "z", 20, ("Ex", "Ez"), freqs)  # runs as the oracle's monitor
    n = 110
    sim.run(n)
    ora.run(n)
    err = rel(mon.dft(), host.acc)
    print(f"[dft-oracle] rel error {err:.3e}", flush=True)
    assert err < 1e-5


def test_plane_dft_reset_and_errors():
    sim = hip().HIPFDTD((16, 16, 16), DL, npml=2)
    P = hip().dft.HIPPlaneDFTMonitor
    with pytest.raises(ValueError):
        P(sim, "w", 3)
    with pytest.raises(ValueError):
        P(sim, "z", 3, ("Bx",))
    with pytest.raises(ValueError):
        P(sim, "z", 3, ("Ex",), np.linspace(1e9, 2e9, 17))
    with pytest.raises(ValueError):
        P(sim, "z", 99, ("Ex",))
    with pytest.raises(ValueError):
        P(sim, "z", 3, ("Ex",), K=7)
    sim64 = hip().HIPFDTD((16, 16, 16), DL, npml=2, dtype=np.float64)
    with pytest.raises(ValueError, match="FP32"):
        P(sim64, "z", 3, ("Ex",))
    m = P(sim64, "z", 3, ("Ex",), [1e10], mode="direct")
    sim64.add_source(lambda f: f.add_source_J("x", waveform(f.t), (7, 8, 7, 8, 7, 8)))
    sim64.run(20)
    assert np.any(m.dft() != 0)
    m.reset()
    assert np.all(m.dft() == 0)
