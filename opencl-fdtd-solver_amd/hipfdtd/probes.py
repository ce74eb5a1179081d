"""Device point probe: running DFT of one field sample per step.

Replaces the reference harness's ``read_point`` + host DFT loop
(``tests/meep_validation/cases.py:179-197``), which synchronises the device
every step, with one small kernel per step that accumulates
``Σ_n F(i,j,k) e^{jω t_{n+1}} dt`` in FP64 on the device.
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib

COMPONENTS = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")


def phase_table(freqs, t0: float, dt: float, n_steps: int) -> np.ndarray:
    """(n_steps, nfreq) complex phases e^{jω t} dt at t = t0 + (n+1) dt, from
    the recurrence phase_{n+1} = phase_n e^{jω dt} used by the monitors
    (reference ``monitors.py:456-459, 571-576``)."""
    om = 2.0 * np.pi * np.atleast_1d(np.asarray(freqs, dtype=np.float64))
    out = np.empty((int(n_steps), om.size), dtype=np.complex128)
    ph = np.exp(1j * om * (t0 + dt)) * dt
    d = np.exp(1j * om * dt)
    for n in range(int(n_steps)):
        out[n] = ph
        ph = ph * d
    return out


class HIPPointProbe:
    """DFT of component ``comp`` at cell (i, j, k) for ``freqs`` (monitor callback)."""

    def __init__(self, fdtd, ijk, comp="Ex", freqs=(1e9,)):
        self.fdtd = fdtd
        self.ijk = tuple(int(v) for v in ijk)
        self.comp = comp
        self.freqs = np.atleast_1d(np.asarray(freqs, dtype=np.float64))
        self.omegas = 2.0 * np.pi * self.freqs
        self._lib = fdtd._lib
        h = C.c_void_p()
        _lib.check(self._lib.hfdtd_probe_create(fdtd._handle, *self.ijk, COMPONENTS.index(comp),
                                                len(self.freqs), C.byref(h)), "probe_create")  # fmt: skip
        self._handle = h
        self._phases = None
        self._dphase = np.exp(1j * self.omegas * fdtd.dt)
        self._buf = np.zeros(2 * len(self.freqs), dtype=np.float64)
        fdtd.add_monitor(self)

    # its per-step call reads the field as the engine holds it, so a
    # row-padded run keeps going while it samples (HIPFDTD.run)
    padded_ok = True

    def __call__(self, fdtd):
        if self._phases is None:
            self._phases = np.exp(1j * self.omegas * fdtd.t) * fdtd.dt
        else:
            self._phases = self._phases * self._dphase
        self._buf[0::2] = self._phases.real
        self._buf[1::2] = self._phases.imag
        _lib.check(self._lib.hfdtd_probe_accumulate(
            self._handle, self._buf.ctypes.data_as(C.POINTER(C.c_double))), "probe")  # fmt: skip

    def watch_boxes(self):
        """The probed cell, for the pair path."""
        i, j, k = self.ijk
        return [(i, i + 1, j, j + 1, k, k + 1, 1 << COMPONENTS.index(self.comp))]

    def dft(self) -> np.ndarray:
        out = np.empty(2 * len(self.freqs), dtype=np.float64)
        _lib.check(self._lib.hfdtd_probe_read(self._handle,
                                              out.ctypes.data_as(C.POINTER(C.c_double))))  # fmt: skip
        return out[0::2] + 1j * out[1::2]

    def _release(self):
        if getattr(self, "_handle", None) is not None:
            self._lib.hfdtd_probe_destroy(self._handle)
            self._handle = None
