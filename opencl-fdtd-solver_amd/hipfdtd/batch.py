"""Batched sweep: many independent ``HIPFDTD`` grids advanced together.

No counterpart in the reference (which runs one grid per ``OpenCLFDTD``).  A
parameter or design sweep of uneven grid sizes becomes one persistent launch
per step over a work list of (grid, tile) pairs, with the K-step loop
captured once as a HIP graph and replayed.  Each grid may carry one current
or soft source with a precomputed amplitude table (the waveform evaluated at
``t_n + dt/2`` on the host, exactly as a source callback would) and one
E-field point probe with a precomputed phase table.
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .probes import phase_table


class BatchedFDTD:
    """Advance ``sims`` (``HIPFDTD`` objects with the same dl/dtype/device)."""

    def __init__(self, sims):
        self.sims = list(sims)
        if not self.sims:
            raise ValueError("BatchedFDTD needs at least one simulation")
        self._lib = self.sims[0]._lib
        arr = (C.c_void_p * len(self.sims))(*[s._handle for s in self.sims])
        h = C.c_void_p()
        _lib.check(self._lib.hfdtd_batch_create(arr, len(self.sims), C.byref(h)), "batch_create")
        self._handle = h
        self._n_table = 0
        self._step = 0
        self._nfreq = 0
        self.path = "fused"

    def close(self):
        if self._handle is not None:
            self._lib.hfdtd_batch_destroy(self._handle)
            self._handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_sources(self, sources, waveforms, n_steps: int):
        """``sources[s]`` is ``(component, box)`` for a current source
        ``E_c += -dt/(ε0 εr) J`` on the half-open box, or ``(component, box,
        'soft')``; ``waveforms[s](t)`` gives J (or the soft amplitude) and is
        sampled at ``t_n + dt/2`` for the next ``n_steps`` steps."""
        n = len(self.sims)
        if len(sources) != n or len(waveforms) != n:
            raise ValueError("one source and one waveform per simulation")
        descs = (_lib.Source * n)()
        amp = np.empty((n, int(n_steps)), dtype=np.float64)
        for s, (src, wf, sim) in enumerate(zip(sources, waveforms, self.sims)):
            comp, box = src[0], src[1]
            mode = _lib.SRC_SOFT if len(src) > 2 and src[2] == "soft" else _lib.SRC_CURRENT
            c = "xyz".index(comp) if isinstance(comp, str) else int(comp)
            i0, i1, j0, j1, k0, k1 = (int(v) for v in box)
            descs[s] = _lib.Source(c, mode, i0, i1, j0, j1, k0, k1, 0, 0.8, 0.0)
            t = sim.t
            for q in range(int(n_steps)):
                amp[s, q] = float(wf(t + 0.5 * sim.dt))
                t += sim.dt
        _lib.check(self._lib.hfdtd_batch_set_sources(
            self._handle, descs, amp.ctypes.data_as(C.POINTER(C.c_double)), int(n_steps)),
            "batch_set_sources")  # fmt: skip
        self._n_table = int(n_steps)
        self._step = 0

    def set_probes(self, cells, comp, freqs_per_sim, n_steps: int):
        """One E-field probe per simulation: ``cells[s] = (i, j, k)`` (or None),
        ``freqs_per_sim[s]`` its DFT frequencies (same count for every sim)."""
        n = len(self.sims)
        nf = len(freqs_per_sim[0])
        ijk = np.full((n, 3), -1, dtype=np.int32)
        ph = np.zeros((n, int(n_steps), nf, 2), dtype=np.float64)
        for s, (cell, fr, sim) in enumerate(zip(cells, freqs_per_sim, self.sims)):
            if cell is None:
                continue
            ijk[s] = cell
            tab = phase_table(fr, sim.t, sim.dt, n_steps)
            ph[s, :, :, 0] = tab.real
            ph[s, :, :, 1] = tab.imag
        c = "xyz".index(comp[-1]) if isinstance(comp, str) else int(comp)
        _lib.check(self._lib.hfdtd_batch_set_probes(
            self._handle, ijk.ctypes.data_as(C.POINTER(C.c_int)), c, nf,
            ph.ctypes.data_as(C.POINTER(C.c_double)), int(n_steps)), "batch_set_probes")  # fmt: skip
        self._nfreq = nf
        self._n_table = int(n_steps)
        self._step = 0

    def set_launch(self, blocks: int = 0, rows_j: int = 0, planes_i: int = 0):
        _lib.check(self._lib.hfdtd_batch_set_launch(self._handle, int(blocks), int(rows_j),
                                                    int(planes_i)))  # fmt: skip

    def set_path(self, path: str) -> None:
        """'fused' (one persistent fused-tile launch per step, the default),
        'split' (persistent split sweeps over flattened planes: H, sources,
        E, probes; no idle lanes on extents that are not multiples of 64) or
        'march' (one launch per step of march tiles over every grid, sources
        and probes handled inside the sweep)."""
        codes = {"fused": _lib.PATH_FUSED, "split": _lib.PATH_SPLIT, "march": _lib.PATH_MARCH}
        if path not in codes:
            raise ValueError("path must be 'fused', 'split' or 'march'")
        _lib.check(self._lib.hfdtd_batch_set_path(self._handle, codes[path]), "batch_set_path")
        self.path = path

    def set_split(self, runs: int = 0) -> None:
        """Work items of the 'split' path: runs of 256 cells of one plane per
        block (1 to 8; 0 keeps the setting, default 2)."""
        _lib.check(self._lib.hfdtd_batch_set_split(self._handle, int(runs)), "batch_set_split")

    def set_march(self, rows: int = 0, planes: int = 0, run: int = 0, width: int = 0) -> None:
        """Launch shape of the 'march' path: j rows per wavefront (1 or 2),
        i planes per chunk, consecutive work items per XCD run, tile width
        in k (64 or 32 columns; 32-column tiles take fewer tiles where nz
        modulo 64 lies in (0, 32]); 0 keeps a setting."""
        _lib.check(self._lib.hfdtd_batch_set_march(self._handle, int(rows), int(planes), int(run),
                                                   int(width)), "batch_set_march")  # fmt: skip

    def set_march_padding(self, on: bool) -> None:
        """Row padding of the 'march' path (on by default).  That path runs
        every engine on a copy in a batch-owned arena, filled before each
        run and written back after it; with padding the copy's rows are
        padded to whole 128-byte lines where nz does not fill them
        (identical results)."""
        _lib.check(self._lib.hfdtd_batch_set_march_padding(self._handle, int(bool(on))),
                   "batch_set_march_padding")  # fmt: skip

    def run(self, n_steps: int, graph_steps: int = 0):
        """Advance every simulation ``n_steps`` steps (one launch per step, or
        graph replays of ``graph_steps`` steps)."""
        _lib.check(self._lib.hfdtd_batch_run(self._handle, int(n_steps), int(graph_steps)),
                   "batch_run")  # fmt: skip
        for sim in self.sims:
            for _ in range(int(n_steps)):
                sim.t += sim.dt
            sim.step_num += int(n_steps)
        self._step += int(n_steps)

    def probe_dfts(self) -> np.ndarray:
        """(n_sims, nfreq) complex128 probe DFTs."""
        n = len(self.sims)
        out = np.empty(n * self._nfreq * 2, dtype=np.float64)
        _lib.check(self._lib.hfdtd_batch_read_probes(self._handle,
                                                     out.ctypes.data_as(C.POINTER(C.c_double))))  # fmt: skip
        out = out.reshape(n, self._nfreq, 2)
        return out[..., 0] + 1j * out[..., 1]
