"""``HIPFDTD`` — drop-in for ``OpenCLFDTD`` (reference ``engine.py:85-789``).

Same constructor, attributes, methods, callback protocol and exceptions; the
work is done by hand-written HIP kernels in ``libhipfdtd.so`` through the
C ABI of ``include/hipfdtd.h``.  All device memory is owned by the library.

Step execution.  The reference runs ``update_H``, then the source callbacks
(at ``t + dt/2``), then ``update_E``, then the monitors
(``engine.py:724-737``).  ``HIPFDTD`` keeps that order on every path:

* ``"split"``: ``update_H``, the source callbacks (each ``add_source_*``
  launches at once), ``update_E``; one thread per cell over the flattened
  (j, k) planes, in place;
* ``"fused"``: during the source phase, ``add_source_*`` calls are recorded
  instead of launched; one LDS-staged sweep advances H and E everywhere,
  then a small fix-up kernel recomputes E on the source cells with the
  recorded sources added first, which is the reference's order cell by
  cell;
* ``"march"``: one register-pipelined sweep advances H and E and applies up
  to four recorded sources in its E half (more go through the fix-up
  kernel); bit-identical to ``"fused"``;
* ``"pair"`` (opt-in, FP32): ``run`` advances two steps per launch,
  bit-identical to two march steps; monitors see the state between the two
  steps through a shadow copy of the cells they name.

On the one-launch paths a callback that touches field data (``Ex``,
``read_point``, a field setter ...) first forces the H half to run on its
own, so it sees exactly the state the reference would show it; the step
then finishes with the split E sweep.

``path=None`` (the default) picks the path measured faster on MI355X: the
march sweep from 200^3 cells up, below that the split sweeps for FP32 and
the fused sweep for FP64.  Paths whose addressing does not reach a large
grid are not offered for it (the fused and pair sweeps need the six-volume
slab below 4 GB and 2 GB; the march sweep takes FP32 grids past that with
one buffer resource per volume).  ``hipfdtd.autotune.tune_engine``
measures the paths and tilings for a given grid.
"""

from __future__ import annotations

import ctypes as C
import logging
import math
import warnings

import numpy as np

from . import _lib
from .constants import C0, EPS0, MU0
from .plugin import SourceMonitorMixin

_LOG = logging.getLogger(__name__)

_MAX_FUSED_SOURCES = 16


class HIPFDTD(SourceMonitorMixin):
    """3-D Yee FDTD solver on one MI355X (or any gfx950 device).

    shape : (Nx, Ny, Nz) cells; dl : cell size [m] (``dx=`` is an alias);
    npml : CPML thickness in cells; dtype : float32 or float64;
    accum_fp64 : N2F DFT / far-field accumulators in double.
    ``ctx``/``queue`` are accepted for signature compatibility and ignored.
    """

    MEMORY_HEADROOM_FRACTION = 0.12
    MEMORY_HEADROOM_BYTES = 512 * 1024 * 1024

    def __init__(self, shape, dl=None, npml=20, dtype=np.float32, ctx=None, queue=None,
                 accum_fp64=False, *, device=0, dx=None, fused=None, path=None):  # fmt: skip
        if dl is None:
            dl = dx
        if dl is None:
            raise TypeError("HIPFDTD needs the cell size dl (or dx=)")
        self.Nx, self.Ny, self.Nz = (int(s) for s in shape)
        self.dl = float(dl)
        self.npml = int(npml)
        dtype = np.dtype(dtype)
        if dtype not in (np.dtype(np.float32), np.dtype(np.float64)):
            raise ValueError(f"HIPFDTD supports float32 or float64 computation; got {dtype!r}")
        self.dtype = dtype
        self.real = np.float32 if dtype == np.dtype(np.float32) else np.float64
        self.complex_dtype = (
            np.dtype(np.complex64) if dtype == np.dtype(np.float32) else np.dtype(np.complex128)
        )
        self.accum_fp64 = bool(accum_fp64) or dtype == np.dtype(np.float64)
        self.accreal = np.float64 if self.accum_fp64 else self.real
        self.accreal_complex_dtype = (
            np.dtype(np.complex128) if self.accum_fp64 else self.complex_dtype
        )
        self.t = 0.0
        self.step_num = 0
        self.dt = 0.99 * self.dl / (C0 * np.sqrt(3.0))
        self.size = self.Nx * self.Ny * self.Nz
        n = self.npml
        self.psi_x_size = (2 * n * self.Ny * self.Nz) if n > 0 else 0
        self.psi_y_size = (self.Nx * 2 * n * self.Nz) if n > 0 else 0
        self.psi_z_size = (self.Nx * self.Ny * 2 * n) if n > 0 else 0
        self.device = int(device)
        # the fused sweep addresses its slabs with 32-bit byte offsets
        item = np.dtype(self.dtype).itemsize
        vstride = (self.Nx * self.Ny * self.Nz + 64 + 63) // 64 * 64
        psi_total = 4 * (self.psi_x_size + self.psi_y_size + self.psi_z_size) + 64
        self._fused_fits = max(6 * vstride, psi_total) * item < 2**32
        self._pair_fits = max(6 * vstride, psi_total) * item < 2**31 and item == 4
        # the march sweep also takes FP32 slabs of 4 GB and more, addressing
        # each volume (below 2 GB) on its own: up to about 810^3 cells
        self._march_fits = self._fused_fits or (item == 4 and max(vstride, psi_total) * item < 2**31)
        self.path = "split"
        self.fused = False
        self.set_path(self._resolve_path(fused, path, dtype, (self.Nx, self.Ny, self.Nz)))
        self._sources = []
        self._monitors = []
        self.lossy = False  # σ or ca volumes set: the sweeps read ca
        self._handle = None
        self._deferring = False
        self._pairing = False
        self._watch_set = None
        self._h_done = False
        self._pending: list[_lib.Source] = []
        self._row_padding = True  # run(): march steps on a row-padded copy where it pays

        self._lib = _lib.load()
        ndev = _lib.device_count()
        if ndev < 1:
            raise RuntimeError("No HIP devices found (HIPFDTD needs a gfx950 GPU)")
        if not 0 <= self.device < ndev:
            raise ValueError(f"device {self.device} out of range ({ndev} devices)")
        total, _, self.device_name = _lib.device_info(self.device)
        self._check_device_memory((self.Nx, self.Ny, self.Nz), self.npml, self.dtype, total)
        h = C.c_void_p()
        _lib.check(
            self._lib.hfdtd_create(self.Nx, self.Ny, self.Nz, self.npml, self.dl, self.dt,
                                   int(self.dtype == np.float64), self.device, C.byref(h)),
            "hfdtd_create",
        )  # fmt: skip
        self._handle = h
        _LOG.info("HIP FDTD solver initialized on device: %s (dtype=%s)", self.device_name,
                  self.dtype.name)  # fmt: skip

    # ── lifetime ──

    def close(self) -> None:
        """Release device memory (also done when the object is collected)."""
        for mon in list(self._monitors):
            rel = getattr(mon, "_release", None)
            if rel is not None:
                rel()
        if self._handle is not None:
            self._lib.hfdtd_destroy(self._handle)
            self._handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ── memory budget (engine.py:320-363) ──

    PATHS = ("split", "fused", "march", "pair")
    # grids of at least this many cells default to the march sweep
    MARCH_DEFAULT_CELLS = 200**3

    @classmethod
    def default_path(cls, dtype, shape=None) -> str:
        """Path chosen when neither ``fused`` nor ``path`` is given, from
        MI355X measurements: the march sweep on grids of 200^3 cells and more
        (256^3 FP32: 264 µs against 303 µs split; FP64 528 against 651 µs
        fused; 512^3 FP32 1.80 against 2.23 ms split), below that the split
        sweeps for FP32 (160^3: 66 µs against 88 µs march) and the LDS-DMA
        fused sweep for FP64.  ``hipfdtd.autotune.tune_engine`` measures all
        of them for a given grid."""
        if shape is not None and int(np.prod([int(v) for v in shape])) >= cls.MARCH_DEFAULT_CELLS:
            return "march"
        return "fused" if np.dtype(dtype) == np.dtype(np.float64) else "split"

    @classmethod
    def default_fused(cls, dtype, shape=None) -> bool:
        """Whether the default path is a one-launch (ping-pong) sweep."""
        return cls.default_path(dtype, shape) != "split"

    @classmethod
    def _resolve_path(cls, fused, path, dtype, shape=None) -> str:
        if path is not None:
            if path not in cls.PATHS:
                raise ValueError(f"path must be one of {cls.PATHS}; got {path!r}")
            return path
        if fused is None:
            return cls.default_path(dtype, shape)
        return "fused" if fused else "split"

    def set_path(self, path) -> None:
        """Step path: 'split' (two sweeps, in place), 'fused' (LDS-DMA tiles),
        'march' (register-pipelined tiles) or 'pair' (``run`` advances two
        steps per launch, single steps take the march sweep); ``True`` /
        ``False`` mean 'fused' / 'split'.  'fused' falls back to 'split' on
        slabs of 4 GB and more, 'march' on FP64 slabs of 4 GB and more or
        FP32 volumes of 2 GB and more, 'pair' to 'march' for FP64 or slabs of
        2 GB and more.  Takes effect at the next step; the state carries
        over."""
        if isinstance(path, (bool, np.bool_)):
            path = "fused" if path else "split"
        if path not in self.PATHS:
            raise ValueError(f"path must be one of {self.PATHS}; got {path!r}")
        if path == "pair" and not self._pair_fits:
            path = "march"
        if path == "march" and not self._march_fits:
            path = "split"
        if path == "fused" and not self._fused_fits:
            path = "split"
        self.path = path
        self.fused = path != "split"

    @staticmethod
    def estimate_device_memory_bytes(shape, npml, dtype=np.float32, fused=None):
        """Device bytes for the 12 volume arrays, the face psi stripes and,
        with ``fused`` (default: whether the path picked for ``dtype`` and
        ``shape`` is a one-launch sweep), the ping-pong copies of fields and
        psi.

        The reference counts 9 volume arrays (``engine.py:325``) and leaves out
        the three ``ca`` arrays it allocates; this estimate counts all twelve.
        """
        nx, ny, nz = (int(s) for s in shape)
        item = np.dtype(dtype).itemsize
        vol = nx * ny * nz * item
        psi = 0
        if npml > 0:
            psi = 4 * (2 * npml * ny * nz + nx * 2 * npml * nz + nx * ny * 2 * npml) * item
        total = 12 * vol + psi
        if fused is None:
            fused = HIPFDTD.default_fused(dtype, shape)
        if fused:
            total += 6 * vol + psi
        return total

    @classmethod
    def device_memory_budget_bytes(cls, device):
        """Usable device bytes after headroom; ``device`` is an ordinal or an
        object with ``global_mem_size``."""
        if hasattr(device, "global_mem_size"):
            total = int(device.global_mem_size)
        else:
            total = _lib.device_info(int(device))[0]
        reserve = max(int(total * cls.MEMORY_HEADROOM_FRACTION), int(cls.MEMORY_HEADROOM_BYTES))
        return max(0, total - reserve)

    def _check_device_memory(self, shape, npml, dtype, total):
        needed = self.estimate_device_memory_bytes(shape, npml, dtype, fused=self.fused)
        reserve = max(int(total * self.MEMORY_HEADROOM_FRACTION), int(self.MEMORY_HEADROOM_BYTES))
        budget = max(0, total - reserve)
        if needed > budget:
            raise MemoryError(
                f"Model needs ~{needed / 1024**3:.2f} GB device memory, but "
                f"{self.device_name} only has ~{budget / 1024**3:.2f} GB usable "
                f"({total / 1024**3:.2f} GB total minus {reserve / 1024**3:.2f} GB headroom). "
                "Reduce the grid or npml."
            )

    def device_memory_bytes(self) -> int:
        """Bytes currently held on the device by this engine."""
        out = C.c_uint64(0)
        _lib.check(self._lib.hfdtd_memory_bytes(self._handle, C.byref(out)))
        return int(out.value)

    # ── materials (engine.py:212-254) ──

    def set_epsilon(self, eps_array, sigma_array=None):
        """Cell-centred εr (and optional σ [S/m]); edge coefficients are built
        on the device (``ce = dt/(ε0 εe)``, or lossy ``ca``/``cb``)."""
        expected = (self.Nx, self.Ny, self.Nz)
        if tuple(eps_array.shape) != expected:
            raise ValueError(
                f"Epsilon shape mismatch: expected {expected}, got {tuple(eps_array.shape)}"
            )
        if sigma_array is not None and tuple(sigma_array.shape) != expected:
            raise ValueError(
                f"Sigma shape mismatch: expected {expected}, got {tuple(sigma_array.shape)}"
            )
        self._materialize()
        src64 = np.asarray(eps_array).dtype == np.float64 or (
            sigma_array is not None and np.asarray(sigma_array).dtype == np.float64
        )
        st = np.float64 if src64 else np.float32
        eps = np.ascontiguousarray(eps_array, dtype=st)
        sig = None if sigma_array is None else np.ascontiguousarray(sigma_array, dtype=st)
        _lib.check(
            self._lib.hfdtd_set_materials(self._handle, eps.ctypes.data_as(C.c_void_p),
                                          None if sig is None else sig.ctypes.data_as(C.c_void_p),
                                          int(src64)),
            "set_epsilon",
        )  # fmt: skip
        self.lossy = sig is not None

    # ── sources (engine.py:448-538) ──

    def _sheet(self, i0, i1, j0, j1):
        return (
            0 if i0 is None else int(i0),
            self.Nx if i1 is None else int(i1),
            0 if j0 is None else int(j0),
            self.Ny if j1 is None else int(j1),
        )

    def _submit(self, src: _lib.Source) -> None:
        if self._deferring and not self._h_done:
            if len(self._pending) < _MAX_FUSED_SOURCES:
                self._pending.append(src)
                return
            self._materialize()
        arr = (_lib.Source * 1)(src)
        _lib.check(self._lib.hfdtd_add_sources(self._handle, arr, 1), "add_source")

    def add_source_Ex(self, z_src, amp, i0=None, i1=None, j0=None, j1=None):
        """Soft-add ``amp`` onto Ex on the sheet z = z_src (legacy inject)."""
        warnings.warn(
            "add_source_Ex is a legacy Ex soft-add; prefer add_source_Jx for SI current density",
            DeprecationWarning,
            stacklevel=2,
        )
        i0, i1, j0, j1 = self._sheet(i0, i1, j0, j1)
        z = int(z_src)
        self._submit(_lib.Source(0, _lib.SRC_SOFT, i0, i1, j0, j1, z, z + 1, 0, 0.0, float(amp)))

    def add_source_Jx(self, z_src, Jx, i0=None, i1=None, j0=None, j1=None, *,
                      rim_taper=False, rim_edge=0.8, rim_renorm=True):  # fmt: skip
        """Inject SI current density Jx [A/m²] on a constant-z Ex sheet:
        ``Ex += -dt/(ε0 εr) Jx`` with optional rim taper and renormalisation."""
        i0, i1, j0, j1 = self._sheet(i0, i1, j0, j1)
        jx = float(Jx)
        re = float(rim_edge)
        if rim_taper and rim_renorm:
            nx_s, ny_s = max(0, i1 - i0), max(0, j1 - j0)
            if nx_s >= 2 and ny_s >= 2:
                ni, nj = nx_s - 2, ny_s - 2
                wsum = ni * nj + re * (2 * ni + 2 * nj) + (re * re) * 4
                jx *= (nx_s * ny_s) / wsum
        z = int(z_src)
        taper = 3 if rim_taper else 0
        self._submit(_lib.Source(0, _lib.SRC_CURRENT, i0, i1, j0, j1, z, z + 1, taper, re, jx))

    def add_source_J(self, component, J, box, *, rim_taper=0, rim_edge=0.8):
        """Current density on component 'x'|'y'|'z' over the half-open box
        ``(i0, i1, j0, j1, k0, k1)``: ``E_c += -dt/(ε0 εr) J``.  A point
        dipole is a 1×1×1 box.  ``rim_taper`` is a bitmask of tapered axes."""
        comp = "xyz".index(component) if isinstance(component, str) else int(component)
        i0, i1, j0, j1, k0, k1 = (int(v) for v in box)
        self._submit(_lib.Source(comp, _lib.SRC_CURRENT, i0, i1, j0, j1, k0, k1,
                                 int(rim_taper), float(rim_edge), float(J)))  # fmt: skip

    # ── stepping (engine.py:724-744) ──

    def _materialize(self) -> None:
        """Inside the source phase: run the H half now and flush recorded sources."""
        if self._pairing:
            raise RuntimeError("a source callback read the fields while the pair path evaluated "
                               "the sources of two steps; use set_path('march') for such "
                               "callbacks")
        if self._deferring and not self._h_done:
            _lib.check(self._lib.hfdtd_update_H(self._handle), "update_H")
            self._h_done = True
            if self._pending:
                arr = (_lib.Source * len(self._pending))(*self._pending)
                _lib.check(self._lib.hfdtd_add_sources(self._handle, arr, len(self._pending)))
                self._pending = []

    def step(self):
        """One Yee step: H, sources at t+dt/2, E, then monitors."""
        self._pending = []
        self._h_done = not self.fused
        if self._h_done:
            _lib.check(self._lib.hfdtd_update_H(self._handle), "update_H")
        t_int = self.t
        self.t = t_int + 0.5 * self.dt
        self._deferring = True
        try:
            for src in self._sources:
                src(self)
        finally:
            self._deferring = False
            self.t = t_int
        if self._h_done:
            _lib.check(self._lib.hfdtd_update_E(self._handle), "update_E")
        else:
            n = len(self._pending)
            arr = (_lib.Source * max(n, 1))(*self._pending)
            _lib.check(self._lib.hfdtd_step(self._handle, arr, n, _lib.PATH_CODES[self.path]),
                       "step")  # fmt: skip
        self._pending = []
        self.t += self.dt
        self.step_num += 1
        for mon in self._monitors:
            mon(self)

    def set_row_padding(self, on: bool) -> None:
        """Row-padded march runs in ``run`` (on by default; extension).  Where
        nz does not fill whole 128-byte lines, ``run`` on the march path steps
        a copy of the state whose rows are padded to whole lines
        (``hfdtd_padded_begin``), written back when the run ends or anything
        else touches the engine; results are identical.  Huygens-box
        monitors and point probes sample the copy; with any other monitor
        the run keeps the engine's own layout."""
        self._row_padding = bool(on)

    def _padded_begin(self) -> bool:
        """Start a row-padded run where it applies; True when one is live."""
        if not (self._row_padding and self.path == "march"
                and all(getattr(m, "padded_ok", False) for m in self._monitors)
                and self.Nz % (128 // self.dtype.itemsize) != 0):  # fmt: skip
            return False
        if self._lib.hfdtd_padded_begin(self._handle) != _lib.OK:
            return False  # no room for the copy, or past the sweep's limits: run as is
        live = C.c_int(0)
        _lib.check(self._lib.hfdtd_padded_live(self._handle, C.byref(live)), "padded_live")
        return bool(live.value)

    def run(self, n_steps, progress_every=0):
        """Run ``n_steps`` steps (on the 'pair' path two per launch where it
        can; on the march path, on a row-padded copy where nz does not fill
        whole cache lines and every monitor samples such a copy)."""
        n = int(n_steps)
        padded = n > 1 and self._padded_begin()
        try:
            self._run(n, n_steps, progress_every)
        finally:
            if padded:
                _lib.check(self._lib.hfdtd_padded_end(self._handle), "padded_end")

    def _run(self, n, n_steps, progress_every):
        i = 0
        while i < n:
            if self.path == "pair" and n - i >= 2 and self._step_pair():
                i += 2
                if progress_every and (i - 2) // progress_every != i // progress_every:
                    _LOG.info("  step %d/%d  t=%.3e s", i, n_steps, self.t)
                continue
            self.step()
            if progress_every and i % progress_every == 0:
                _LOG.info("  step %d/%d  t=%.3e s", i, n_steps, self.t)
            i += 1

    def run_device(self, n_steps, source=None, probe=None, *, graph_steps=100):
        """Advance ``n_steps`` steps with a device-resident source and probe,
        the step loop captured once as a HIP graph of ``graph_steps`` steps
        and replayed (extension: the reference evaluates its sources and
        monitors in Python every step).  ``source`` is ``(component, box,
        waveform)`` or ``(component, box, waveform, 'soft')``: a current (or
        soft) source on the half-open box whose ``waveform(t)`` is sampled
        on the host once per step, at t + dt/2, into a table (no rim taper).
        ``probe`` is ``((i, j, k), 'Ex' | 'Ey' | 'Ez', freqs)``; returns its
        DFT Σ E e^{jωt} dt over these steps (complex128, one per frequency),
        or None without a probe.  The results equal ``run`` with the same
        source as a callback and a ``HIPPointProbe``, bit for bit: a split
        engine runs batched split sweeps, a march, fused or pair engine
        batched march tiles (the march sweep's arithmetic, which the fused
        and pair sweeps share).  An engine too large for batched march tiles
        (field slab of 4 GB or more, psi slab of 2 GB or more) runs batched
        split sweeps, equal to its own path to rounding.  Batched march
        tiles run on a copy of the engine (state, psi, coefficients) made
        for the call and written back at its end, which needs that much
        free device memory.  Registered source and monitor callbacks cannot
        run inside a graph: the call refuses an engine that has any."""
        if self._sources or self._monitors:
            raise ValueError("run_device runs no callbacks: clear_sources() / clear_monitors() "
                             "first, or use run()")
        from .batch import BatchedFDTD

        n = int(n_steps)
        if n <= 0:
            return None
        item = np.dtype(self.dtype).itemsize
        psi_bytes = (4 * (self.psi_x_size + self.psi_y_size + self.psi_z_size) + 64) * item
        march = self.path != "split" and self._fused_fits and psi_bytes < 2**31
        batch = BatchedFDTD([self])
        try:
            batch.set_path("march" if march else "split")
            # longer work items on larger planes (fewer block set-ups per cell)
            batch.set_split(4 if self.Ny * self.Nz >= 8192 else 2)
            if source is not None:
                comp, box, wave = source[0], source[1], source[2]
                spec = (comp, box, "soft") if len(source) > 3 and source[3] == "soft" else (comp, box)
                batch.set_sources([spec], [wave], n)
            if probe is not None:
                cell, comp, freqs = probe
                batch.set_probes([tuple(cell)], comp, [list(freqs)], n)
            batch.run(n, graph_steps if graph_steps and n >= graph_steps else 0)
            return batch.probe_dfts()[0] if probe is not None else None
        finally:
            batch.close()

    def _pair_watch(self):
        """Cells the monitors read, as boxes (i0, i1, j0, j1, k0, k1[, mask])
        with an optional mask of the components read (bit c for Ex .. Hz,
        0 = all), or None when a monitor does not name them
        (``watch_boxes``) or there are more boxes than the pair sweep
        records."""
        boxes = []
        for mon in self._monitors:
            fn = getattr(mon, "watch_boxes", None)
            if fn is None:
                return None
            boxes.extend(tuple(int(v) for v in (tuple(b) + (0,))[:7]) for b in fn())
        if len(boxes) > _lib.PAIR_MAX_WATCH:
            return None
        return tuple(boxes)

    def _single(self, pending) -> None:
        """One step on the march sweep with an already evaluated source list."""
        n = len(pending)
        arr = (_lib.Source * max(n, 1))(*pending)
        _lib.check(self._lib.hfdtd_step(self._handle, arr, n, _lib.PATH_MARCH), "step")
        self.t += self.dt
        self.step_num += 1
        for mon in self._monitors:
            mon(self)

    def _step_pair(self) -> bool:
        """Two steps in one launch of the pair sweep: the source callbacks of
        both steps are evaluated first (at t + dt/2 and t + 3dt/2), the
        monitors then run after each step, for the first one on the state
        between the two steps, which the sweep wrote for the cells they
        watch.  False (nothing done) when a monitor cannot name its cells."""
        watch = self._pair_watch()
        if watch is None:
            return False
        lists = []
        t0, n0 = self.t, self.step_num
        # both steps' sources are recorded, never launched: whatever path ran
        # the previous step, no H half of this launch has run yet
        self._h_done = False
        try:
            self._deferring = self._pairing = True
            t_s = t0
            for s in range(2):
                self._pending = []
                # the callbacks of step n0 + s see the time and step count
                # step() shows them (t + dt/2, summed as step() sums it, and
                # step_num n0 + s)
                self.t = t_s + 0.5 * self.dt
                self.step_num = n0 + s
                t_s = t_s + self.dt
                for src in self._sources:
                    src(self)
                lists.append(self._pending)
        finally:
            self._deferring = self._pairing = False
            self._pending = []
            self.t = t0
            self.step_num = n0
        if max(len(v) for v in lists) > _lib.PAIR_MAX_SOURCES:
            for v in lists:
                self._single(v)
            return True
        if watch != self._watch_set:
            flat = [v for b in watch for v in b]
            arr = (C.c_int * max(len(flat), 1))(*flat)
            _lib.check(self._lib.hfdtd_set_watch(self._handle, arr, len(watch)), "set_watch")
            self._watch_set = watch
        a = (_lib.Source * max(len(lists[0]), 1))(*lists[0])
        b = (_lib.Source * max(len(lists[1]), 1))(*lists[1])
        _lib.check(self._lib.hfdtd_step_pair(self._handle, a, len(lists[0]), b, len(lists[1])),
                   "step_pair")  # fmt: skip
        self.t = t0 + self.dt
        self.step_num += 1
        if self._monitors:
            view = len(watch) > 0
            if view:
                _lib.check(self._lib.hfdtd_set_view(self._handle, 1), "set_view")
            try:
                for mon in self._monitors:
                    mon(self)
            finally:
                if view:
                    _lib.check(self._lib.hfdtd_set_view(self._handle, 0), "set_view")
        self.t = (t0 + self.dt) + self.dt
        self.step_num += 1
        for mon in self._monitors:
            mon(self)
        return True

    def run_until_converged(self, monitor, tol=1e-3, check_every=100, max_steps=100_000,
                            min_steps=0, freq_index=0):  # fmt: skip
        """Step until a DFT monitor's sums settle.

        Every ``check_every`` steps the monitor's relative change against its
        last snapshot, ``dft_relative_change`` (reference ``monitors.py:498-569``,
        reduced on the device: one scalar comes back per check), is recorded
        and a new snapshot is taken; the run stops at the first check with a
        change below ``tol`` once ``min_steps`` steps are done, or after
        ``max_steps``.  Between checks nothing synchronises with the host.
        Returns (steps run, [(step, relative change), ...]).
        """
        if check_every < 1:
            raise ValueError("check_every must be >= 1")
        history = []
        done = 0
        monitor.snapshot_dft()
        while done < max_steps:
            n = min(int(check_every), int(max_steps) - done)
            self.run(n)
            done += n
            rel = monitor.dft_relative_change(freq_index)
            monitor.snapshot_dft()
            history.append((self.step_num, rel))
            if done >= min_steps and rel < tol:
                break
        return done, history

    def synchronize(self) -> None:
        """Wait for all queued device work (``queue.finish()``)."""
        _lib.check(self._lib.hfdtd_synchronize(self._handle), "synchronize")

    def reset_fields(self) -> None:
        """Zero E, H and psi; keep materials; t = 0."""
        _lib.check(self._lib.hfdtd_reset_fields(self._handle), "reset_fields")
        self.t = 0.0
        self.step_num = 0

    def set_march_tiling(self, rows_per_wave: int = 0, planes_i: int = 0, run_blocks: int = 0,
                         nontemporal: int = -2) -> None:  # fmt: skip
        """Launch shape of the march sweep: j rows per wavefront (1 or 2;
        until set, one on grids with a material index and two otherwise),
        planes of i per block (-1 = from the grid), consecutive blocks per XCD
        run, non-temporal streams (1 / 0 / -1 = when the slabs exceed the
        last-level cache); 0 (-2 for ``nontemporal``) keeps a setting."""
        _lib.check(self._lib.hfdtd_set_march_tiling(self._handle, int(rows_per_wave),
                                                    int(planes_i), int(run_blocks),
                                                    int(nontemporal)), "set_march_tiling")  # fmt: skip

    def set_march_width(self, width: int = 0) -> None:
        """March-sweep tile width: 64 (default) or 32 columns of k, with
        64 / width rows of j per wavefront slot (FP32 grids with slabs below
        4 GB; others keep 64).  32-column tiles take fewer tiles where nz
        modulo 64 lies in (0, 32].  The results do not depend on it."""
        _lib.check(self._lib.hfdtd_set_march_width(self._handle, int(width)), "set_march_width")

    def set_march_addressing(self, per_volume: bool) -> None:
        """March sweep addressing: one buffer resource per volume on any FP32
        grid (True), or only where the slabs reach 4 GB (False, default);
        the results are the same."""
        _lib.check(self._lib.hfdtd_set_march_addressing(self._handle, int(bool(per_volume))),
                   "set_march_addressing")  # fmt: skip

    def set_overlap(self, on: bool) -> None:
        """True (default): a K-buffered near-to-far monitor's per-step gather
        is held and run by the next march sweep's launch, in blocks after the
        sweep's own (one launch fewer per step); False: each gather is
        launched at once.  The results are the same."""
        _lib.check(self._lib.hfdtd_set_overlap(self._handle, int(bool(on))), "set_overlap")

    def material_count(self) -> int:
        """Materials of the march sweep's material index (0: the sweep reads
        the ce volumes).  Lossless FP32 grids with at most 48 distinct εr
        values get one at ``set_epsilon``."""
        n = C.c_int(0)
        _lib.check(self._lib.hfdtd_material_count(self._handle, C.byref(n)), "material_count")
        return int(n.value)

    CE_SOURCES = ("volumes", "index", "eps")

    def ce_source(self) -> str:
        """Where the march sweep takes ce from: ``'volumes'`` (the three ce
        volumes), ``'index'`` (the material index) or ``'eps'`` (the eps_r
        volume of a lossless FP32 grid without an index, ce recomputed per
        edge exactly as the coefficients were made)."""
        n = C.c_int(0)
        _lib.check(self._lib.hfdtd_ce_source(self._handle, C.byref(n)), "ce_source")
        return self.CE_SOURCES[n.value]

    def set_pair_planes(self, planes_i: int = 0) -> None:
        """Planes of i per block of the pair sweep (0 = 24)."""
        _lib.check(self._lib.hfdtd_set_pair_planes(self._handle, int(planes_i)), "set_pair_planes")

    def set_fused_tiling(self, rows_j: int = 0, planes_i: int = 0, xcd_groups: int = 0) -> None:
        """Launch shape of the fused sweep: rows of j and planes of i per block,
        and the XCD count the work order is grouped for (0 keeps a value;
        planes_i=-1 returns to the grid-derived plane count)."""
        _lib.check(self._lib.hfdtd_set_fused_tiling(self._handle, int(rows_j), int(planes_i),
                                                    int(xcd_groups)))  # fmt: skip

    def set_split_schedule(self, run_blocks: int = 0) -> None:
        """Work order of the split sweeps: runs of ``run_blocks`` consecutive
        256-cell blocks per XCD (1 = hardware round robin, 0 keeps)."""
        _lib.check(self._lib.hfdtd_set_split_schedule(self._handle, int(run_blocks)))

    def kernel_timing(self, enable: bool, every: int = 1) -> None:
        """Start (True) or stop (False) HIP-event timing of the sweep kernels:
        every ``every``-th sweep launch is bracketed by a pair of events (an
        odd ``every`` samples the H and E sweeps of the split path alike)."""
        if int(every) < 1:
            raise ValueError("every must be at least 1")
        _lib.check(self._lib.hfdtd_timing(self._handle, int(every) if enable else 0), "timing")

    def kernel_timing_read(self):
        """(total milliseconds, launches) of the sweep launches timed so far."""
        ms = C.c_double(0.0)
        n = C.c_int(0)
        _lib.check(self._lib.hfdtd_timing_read(self._handle, C.byref(ms), C.byref(n)), "timing")
        return float(ms.value), int(n.value)

    def kernel_timing_launches(self) -> int:
        """Sweep launches since timing was enabled, timed or not."""
        n = C.c_longlong(0)
        _lib.check(self._lib.hfdtd_timing_launches(self._handle, C.byref(n)), "timing")
        return int(n.value)

    # ── field access (engine.py:747-789) ──

    def _read(self, name: str) -> np.ndarray:
        self._materialize()
        out = np.empty((self.Nx, self.Ny, self.Nz), dtype=self.dtype)
        _lib.check(self._lib.hfdtd_read_volume(self._handle, _lib.VOLUME_IDS[name],
                                               out.ctypes.data_as(C.c_void_p)), "read")  # fmt: skip
        return out

    def _write(self, name: str, value) -> None:
        self._materialize()
        arr = np.ascontiguousarray(np.broadcast_to(np.asarray(value, dtype=self.dtype),
                                                   (self.Nx, self.Ny, self.Nz)))  # fmt: skip
        _lib.check(self._lib.hfdtd_write_volume(self._handle, _lib.VOLUME_IDS[name],
                                                arr.ctypes.data_as(C.c_void_p)), "write")  # fmt: skip

    def read_volume(self, name: str) -> np.ndarray:
        """Host copy of a field or coefficient volume (``'Ex'``.. ``'ca_z'``)."""
        if name not in _lib.VOLUME_IDS:
            raise ValueError(f"unknown volume {name!r}")
        return self._read(name)

    def write_volume(self, name: str, value) -> None:
        """Upload a field or coefficient volume."""
        if name not in _lib.VOLUME_IDS:
            raise ValueError(f"unknown volume {name!r}")
        self._write(name, value)
        if name.startswith("ca_"):
            self.lossy = True

    def read_point(self, field_name, i, j, k):
        """One value of a field / coefficient volume."""
        if field_name not in _lib.VOLUME_IDS:
            raise AttributeError(f"no buffer {field_name!r}")
        self._materialize()
        idx = int(i) * self.Ny * self.Nz + int(j) * self.Nz + int(k)
        out = C.c_double(0.0)
        _lib.check(self._lib.hfdtd_read_point(self._handle, _lib.VOLUME_IDS[field_name], idx,
                                              C.byref(out)), "read_point")  # fmt: skip
        return float(out.value)

    Ex = property(lambda s: s._read("Ex"), lambda s, v: s._write("Ex", v))
    Ey = property(lambda s: s._read("Ey"), lambda s, v: s._write("Ey", v))
    Ez = property(lambda s: s._read("Ez"), lambda s, v: s._write("Ez", v))
    Hx = property(lambda s: s._read("Hx"), lambda s, v: s._write("Hx", v))
    Hy = property(lambda s: s._read("Hy"), lambda s, v: s._write("Hy", v))
    Hz = property(lambda s: s._read("Hz"), lambda s, v: s._write("Hz", v))

    def fields(self) -> dict:
        return {n: self._read(n) for n in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")}


def courant_dt(dl: float) -> float:
    """The engine's time step 0.99·dl/(c0·√3) (engine.py:138)."""
    return 0.99 * float(dl) / (C0 * math.sqrt(3.0))


__all__ = ["HIPFDTD", "courant_dt", "EPS0", "MU0"]
