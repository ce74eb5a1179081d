"""Multi-frequency DFT monitor over an axis-aligned plane.

The reference accumulates plane DFTs on the host from field reads, one
synchronising copy per step (``tests/meep_validation/cases.py:179-197``:
``Σ_n Ex(t_{n+1}) e^{jω t_{n+1}} dt``).  ``HIPPlaneDFTMonitor`` keeps the sums
on the device.  Two modes:

* ``mode="buffered"`` (FP32 engines): the plane samples of K steps are kept
  in a device buffer and folded into the sums with one FP32 matrix-core
  contraction per K steps; the sums are touched once per block.
* ``mode="direct"``: per-step accumulation, ``accum`` = ``"fp32"``,
  ``"kahan"`` (compensated FP32) or ``"fp64"``.

Phases ``e^{jω t} dt`` follow the monitors' FP64 recurrence
``phase_{n+1} = phase_n e^{jω dt}`` (reference ``monitors.py:456-459,
571-576``).
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib

COMPONENTS = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")
_AXES = {"x": 0, "y": 1, "z": 2}
_ACCUM = {"fp32": _lib.ACC_REAL, "real": _lib.ACC_REAL, "fp64": _lib.ACC_FP64,
          "kahan": _lib.ACC_KAHAN}  # fmt: skip


class HIPPlaneDFTMonitor:
    """DFT of ``components`` on the plane ``axis = index`` at ``freqs`` (≤ 16).

    Registered as a step monitor on ``fdtd``; ``dft()`` returns a complex128
    array (n_components, n0, n1, n_freq) with (n0, n1) = (Ny, Nz), (Nx, Nz)
    or (Nx, Ny) for x-, y- and z-planes.
    """

    def __init__(self, fdtd, axis="z", index=None, components=("Ex",), freqs=(1e9,), *,
                 mode="buffered", K=32, accum="fp32"):  # fmt: skip
        if axis not in _AXES:
            raise ValueError("axis must be 'x', 'y' or 'z'")
        n_axis = (fdtd.Nx, fdtd.Ny, fdtd.Nz)[_AXES[axis]]
        if index is None:
            index = n_axis // 2
        comps = (components,) if isinstance(components, str) else tuple(components)
        if not comps or any(c not in COMPONENTS for c in comps):
            raise ValueError(f"components must be drawn from {COMPONENTS}")
        if mode not in ("buffered", "direct"):
            raise ValueError("mode must be 'buffered' or 'direct'")
        if accum not in _ACCUM:
            raise ValueError(f"accum must be one of {sorted(_ACCUM)}")
        self.fdtd = fdtd
        self.axis, self.index = axis, int(index)
        self.components = tuple(c for c in COMPONENTS if c in comps)
        self.freqs = np.atleast_1d(np.asarray(freqs, dtype=np.float64))
        if not 1 <= self.freqs.size <= 16:
            raise ValueError("1 to 16 frequencies per plane monitor")
        self.omegas = 2.0 * np.pi * self.freqs
        self.mode, self.K, self.accum = mode, int(K), accum
        self._lib = fdtd._lib
        mask = sum(1 << COMPONENTS.index(c) for c in self.components)
        h = C.c_void_p()
        _lib.check(self._lib.hfdtd_dft_plane_create(
            fdtd._handle, _AXES[axis], self.index, mask, self.freqs.size, self.K,
            _lib.DFT_BUFFERED if mode == "buffered" else _lib.DFT_DIRECT, _ACCUM[accum],
            C.byref(h)), "hfdtd_dft_plane_create")  # fmt: skip
        self._handle = h
        shape = [C.c_int() for _ in range(4)]
        _lib.check(self._lib.hfdtd_dft_shape(h, *[C.byref(v) for v in shape]))
        self.shape = tuple(v.value for v in shape)  # ncomp, n0, n1, nfreq
        self._phases = None
        self._dphase = np.exp(1j * self.omegas * fdtd.dt)
        self._buf = np.zeros(2 * self.freqs.size, dtype=np.float64)
        fdtd.add_monitor(self)

    def __call__(self, fdtd):
        if self._phases is None:
            self._phases = np.exp(1j * self.omegas * fdtd.t) * fdtd.dt
        else:
            self._phases = self._phases * self._dphase
        self._buf[0::2] = self._phases.real
        self._buf[1::2] = self._phases.imag
        _lib.check(self._lib.hfdtd_dft_accumulate(
            self._handle, self._buf.ctypes.data_as(C.POINTER(C.c_double))), "dft")  # fmt: skip

    def watch_boxes(self):
        """The recorded plane, for the pair path."""
        box = [0, self.fdtd.Nx, 0, self.fdtd.Ny, 0, self.fdtd.Nz]
        a = _AXES[self.axis]
        box[2 * a], box[2 * a + 1] = self.index, self.index + 1
        mask = sum(1 << COMPONENTS.index(c) for c in self.components)
        return [tuple(box) + (mask,)]

    def dft(self) -> np.ndarray:
        """Sums so far (a partial block is folded in first)."""
        ncomp, n0, n1, nf = self.shape
        out = np.empty(ncomp * n0 * n1 * nf * 2, dtype=np.float64)
        _lib.check(self._lib.hfdtd_dft_read(self._handle,
                                            out.ctypes.data_as(C.POINTER(C.c_double))))  # fmt: skip
        out = out.reshape(ncomp, n0, n1, nf, 2)
        return out[..., 0] + 1j * out[..., 1]

    def component(self, name: str) -> np.ndarray:
        """(n0, n1, n_freq) DFT of one recorded component."""
        return self.dft()[self.components.index(name)]

    def reset(self) -> None:
        _lib.check(self._lib.hfdtd_dft_reset(self._handle))
        self._phases = None

    def _release(self):
        if getattr(self, "_handle", None) is not None:
            self._lib.hfdtd_dft_destroy(self._handle)
            self._handle = None

    def __del__(self):
        try:
            self._release()
        except Exception:
            pass
