"""Adjoint ε-gradients of probe-DFT objectives, for inverse design.

The reference has no differentiation; this adds it beside the engine.  A
:class:`ProbeDFTAdjoint` problem fixes a grid, current sources, probe cells
and frequencies, a run length and a design box.  ``forward(eps_design)`` runs
the HIP engine and returns the probe DFTs

    F_p(ω) = Σ_{n=1..N} E_p(t_n) e^{jω t_n} dt,   t_n = n dt

(the reference harness's point DFT, ``tests/meep_validation/cases.py:179-197``),
and ``backward(grad_F)`` returns dL/dε on every design cell for a real
objective L(F) whose conjugate-Wirtinger gradient (the one torch hands back)
is ``grad_F``.  :func:`probe_dft` wraps both in a ``torch.autograd.Function``.

The reverse sweep is the exact discrete adjoint of the engine's lossless
Yee + CPML step (``libhipfdtd`` ``hfdtd_adjoint_*``): the state after every
``checkpoint`` steps is kept on the device, each segment is recomputed from
its checkpoint (recording E over the design box every step) and then swept
backwards; the last segment's history is still in place after the forward
run, so it is not recomputed.  Device memory: one state copy per segment
plus ``checkpoint + 1`` snapshots of E over the design box.

Materials enter through the edge average ``ce = dt/(ε0 εe)`` of ``set_epsilon``
(``materials.py:28-56``); the library returns G = Σ_n Ẽ^n (E^n − E^{n−1}) per
edge, dL/dεe = −G/εe, and the edge average is transposed onto the cells here.
Only current sources (``E += −ce J``) and lossless grids are supported: both
keep ∂E^n/∂ce = (E^n − E^{n−1})/ce exact.
"""

from __future__ import annotations

import ctypes as C
import math
import time
from typing import Callable, NamedTuple, Sequence

import numpy as np

from . import _lib
from .engine import HIPFDTD

_COMP = {"x": 0, "y": 1, "z": 2, "Ex": 0, "Ey": 1, "Ez": 2, 0: 0, 1: 1, 2: 2}


class CurrentSource(NamedTuple):
    """Current density J(t) [A/m²] on component 'x'|'y'|'z' over the half-open
    box (i0, i1, j0, j1, k0, k1); step n uses J((n + ½) dt)."""

    component: object
    box: Sequence[int]
    waveform: Callable[[float], float]


def _edge_values(eps, c, box):
    """εe of component c on the cells of ``box`` (materials.py:28-42)."""
    i0, i1, j0, j1, k0, k1 = box
    lo = [i0, j0, k0]
    hi = [i1, j1, k1]
    n = eps.shape[c]
    here = eps[i0:i1, j0:j1, k0:k1]
    hi2 = list(hi)
    hi2[c] = min(hi[c] + 1, n)
    lo2 = list(lo)
    lo2[c] = lo[c] + 1
    nxt = eps[lo2[0]:hi2[0], lo2[1]:hi2[1], lo2[2]:hi2[2]]
    out = here.copy()
    m = nxt.shape[c]  # cells whose +1 neighbour exists
    sl = [slice(None)] * 3
    sl[c] = slice(0, m)
    out[tuple(sl)] = 0.5 * (here[tuple(sl)] + nxt)
    return out


class ProbeDFTAdjoint:
    """Forward probe DFTs and their ε-gradient over a design box on one GPU.

    shape, dl, npml, dtype, device: the grid (as ``HIPFDTD``);
    design_box: (i0, i1, j0, j1, k0, k1) half-open; sources: iterable of
    :class:`CurrentSource` (or (component, box, waveform) tuples, at most 16);
    probes: iterable of (component, i, j, k) with component Ex/Ey/Ez;
    freqs [Hz]; steps: run length N; checkpoint: segment length K;
    eps_background: εr outside the design box (default 1); fused: the
    forward path (``HIPFDTD``'s ``fused``, default per dtype and grid size).
    """

    def __init__(self, shape, dl, npml, design_box, sources, probes, freqs, steps, *,
                 eps_background=None, checkpoint=100, dtype=np.float32, device=0,
                 fused=None):  # fmt: skip
        self.shape = tuple(int(s) for s in shape)
        self.steps = int(steps)
        if self.steps < 1:
            raise ValueError("steps must be >= 1")
        self.K = max(1, min(int(checkpoint), self.steps))
        self.S = math.ceil(self.steps / self.K)
        self.design_box = tuple(int(v) for v in design_box)
        if len(self.design_box) != 6:
            raise ValueError("design_box is (i0, i1, j0, j1, k0, k1)")
        self.fdtd = HIPFDTD(self.shape, dl, npml=npml, dtype=dtype, device=device, fused=fused)
        self._lib = self.fdtd._lib
        if eps_background is None:
            eps_background = np.ones(self.shape)
        self.eps_background = np.array(eps_background, dtype=np.float64)
        if self.eps_background.shape != self.shape:
            raise ValueError(f"eps_background must have shape {self.shape}")
        box = (C.c_int * 6)(*self.design_box)
        h = C.c_void_p()
        _lib.check(self._lib.hfdtd_adjoint_create(self.fdtd._handle, box, self.K + 1, self.S,
                                                  C.byref(h)), "hfdtd_adjoint_create")  # fmt: skip
        self._handle = h
        gb = (C.c_int * 6)()
        _lib.check(self._lib.hfdtd_adjoint_grad_box(h, gb))
        self.grad_box = tuple(gb)

        self.probes = [(_COMP[c], int(i), int(j), int(k)) for c, i, j, k in probes]
        if not self.probes:
            raise ValueError("at least one probe")
        self.freqs = np.atleast_1d(np.asarray(freqs, dtype=np.float64))
        self.omegas = 2.0 * np.pi * self.freqs
        ijkc = np.array([[i, j, k, c] for c, i, j, k in self.probes], dtype=np.int32)
        _lib.check(self._lib.hfdtd_adjoint_set_probes(
            h, len(self.probes), ijkc.ctypes.data_as(C.POINTER(C.c_int)), self.freqs.size,
            self.omegas.ctypes.data_as(C.POINTER(C.c_double))), "set_probes")  # fmt: skip

        srcs = [CurrentSource(*s) for s in sources]
        if len(srcs) > 16:
            raise ValueError("at most 16 sources")
        self._src = (_lib.Source * max(len(srcs), 1))()
        for q, s in enumerate(srcs):
            i0, i1, j0, j1, k0, k1 = (int(v) for v in s.box)
            self._src[q] = _lib.Source(_COMP[s.component], _lib.SRC_CURRENT, i0, i1, j0, j1, k0,
                                       k1, 0, 0.8, 0.0)  # fmt: skip
        self._nsrc = len(srcs)
        t = (np.arange(self.steps, dtype=np.float64) + 0.5) * self.fdtd.dt
        self._amp = np.ascontiguousarray(
            np.array([[float(s.waveform(tt)) for s in srcs] for tt in t], dtype=np.float64)
            .reshape(self.steps, max(self._nsrc, 0)))  # fmt: skip
        self._eps = None
        self._fresh = False  # last segment's history matches the last forward run
        self.token = 0
        self.timing = {}

    # ── helpers ──

    def _seg(self, s):
        n0 = s * self.K
        return n0, min(n0 + self.K, self.steps)

    def _run(self, n0, n1, slot0, probe):
        amp = self._amp[n0:n1]
        ptr = amp.ctypes.data_as(C.POINTER(C.c_double)) if self._nsrc else None
        path = _lib.PATH_CODES[self.fdtd.path]
        _lib.check(self._lib.hfdtd_adjoint_forward(self._handle, self._src, self._nsrc, ptr,
                                                   n1 - n0, n0, slot0, int(probe), path),
                   "adjoint forward")  # fmt: skip

    def _sync(self):
        self.fdtd.synchronize()

    def set_design(self, eps_design) -> None:
        eps_design = np.asarray(eps_design, dtype=np.float64)
        i0, i1, j0, j1, k0, k1 = self.design_box
        if eps_design.shape != (i1 - i0, j1 - j0, k1 - k0):
            raise ValueError(f"eps_design must have shape {(i1 - i0, j1 - j0, k1 - k0)}")
        eps = self.eps_background.copy()
        eps[i0:i1, j0:j1, k0:k1] = eps_design
        self.fdtd.set_epsilon(eps)
        self._eps = eps

    # ── forward / backward ──

    def forward(self, eps_design) -> np.ndarray:
        """Probe DFTs (n_probe, n_freq) complex128 for this design."""
        self.set_design(eps_design)
        self._sync()
        t0 = time.perf_counter()
        self.fdtd.reset_fields()
        _lib.check(self._lib.hfdtd_adjoint_probe_reset(self._handle))
        for s in range(self.S):
            n0, n1 = self._seg(s)
            _lib.check(self._lib.hfdtd_adjoint_save(self._handle, s))
            last = s == self.S - 1
            if last:
                _lib.check(self._lib.hfdtd_adjoint_record(self._handle, 0))
            self._run(n0, n1, 0 if last else -1, True)
        out = np.empty(len(self.probes) * self.freqs.size * 2, dtype=np.float64)
        _lib.check(self._lib.hfdtd_adjoint_probe_read(self._handle,
                                                      out.ctypes.data_as(C.POINTER(C.c_double))))  # fmt: skip
        self.timing = {"forward_s": time.perf_counter() - t0}
        self.fdtd.t = self.steps * self.fdtd.dt
        self.fdtd.step_num = self.steps
        self._fresh = True
        self.token += 1
        out = out.reshape(len(self.probes), self.freqs.size, 2)
        return out[..., 0] + 1j * out[..., 1]

    def backward(self, grad_F) -> np.ndarray:
        """dL/dε over the design box for the conjugate-Wirtinger gradient
        ``grad_F`` of L at the last forward run's F (2F for L = Σ|F|²)."""
        if self._eps is None:
            raise RuntimeError("backward() needs a forward() run first")
        gF = np.asarray(grad_F, dtype=np.complex128).reshape(len(self.probes), self.freqs.size)
        coef = np.empty(gF.shape + (2,), dtype=np.float64)
        coef[..., 0] = gF.real
        coef[..., 1] = -gF.imag
        h = self._handle
        _lib.check(self._lib.hfdtd_adjoint_set_objective_grad(
            h, coef.ctypes.data_as(C.POINTER(C.c_double))), "set_objective_grad")  # fmt: skip
        self._sync()
        t0 = time.perf_counter()
        recompute = reverse = 0.0
        _lib.check(self._lib.hfdtd_adjoint_begin(h, self.steps), "adjoint begin")
        for s in range(self.S - 1, -1, -1):
            n0, n1 = self._seg(s)
            if not (s == self.S - 1 and self._fresh):
                ta = time.perf_counter()
                _lib.check(self._lib.hfdtd_adjoint_restore(h, s))
                _lib.check(self._lib.hfdtd_adjoint_record(h, 0))
                self._run(n0, n1, 0, False)
                self._sync()
                recompute += time.perf_counter() - ta
            tb = time.perf_counter()
            _lib.check(self._lib.hfdtd_adjoint_reverse(h, n1, n0, n1 - n0), "adjoint reverse")
            self._sync()
            reverse += time.perf_counter() - tb
        self._fresh = False
        gi0, gi1, gj0, gj1, gk0, gk1 = self.grad_box
        bshape = (gi1 - gi0, gj1 - gj0, gk1 - gk0)
        G = np.empty(3 * int(np.prod(bshape)), dtype=np.float64)
        _lib.check(self._lib.hfdtd_adjoint_read_grad(h, G.ctypes.data_as(C.POINTER(C.c_double))))
        grad = self._cells_from_edges(G.reshape((3,) + bshape))
        self.timing.update(backward_s=time.perf_counter() - t0, recompute_s=recompute,
                           reverse_s=reverse)  # fmt: skip
        return grad

    def value_and_grad(self, eps_design):
        """(J, dJ/dε) for J = Σ_{p,f} |F_pf|²."""
        F = self.forward(eps_design)
        return float(np.sum(np.abs(F) ** 2)), self.backward(2.0 * F)

    def _cells_from_edges(self, G):
        """dL/dε per design cell from G per edge (transposed edge average)."""
        i0, i1, j0, j1, k0, k1 = self.design_box
        lo_d = (i0, j0, k0)
        hi_d = (i1, j1, k1)
        lo_b = self.grad_box[0::2]
        out = np.zeros((i1 - i0, j1 - j0, k1 - k0))
        for c in range(3):
            ge = -G[c] / _edge_values(self._eps, c, self.grad_box)
            own = tuple(slice(lo_d[a] - lo_b[a], hi_d[a] - lo_b[a]) for a in range(3))
            w = np.full(hi_d[c] - lo_d[c], 0.5)
            if hi_d[c] == self.shape[c]:
                w[-1] = 1.0  # the last plane keeps the cell value
            shape = [1, 1, 1]
            shape[c] = -1
            out += ge[own] * w.reshape(shape)
            # the edge below: cell − e_c, present when the cell index is >= 1
            nb = list(own)
            start = lo_d[c] - lo_b[c] - 1
            if start >= 0:
                nb[c] = slice(start, start + hi_d[c] - lo_d[c])
                out += 0.5 * ge[tuple(nb)]
            else:
                nb[c] = slice(0, hi_d[c] - lo_d[c] - 1)
                dst = [slice(None)] * 3
                dst[c] = slice(1, None)
                out[tuple(dst)] += 0.5 * ge[tuple(nb)]
        return out

    def memory_bytes(self) -> int:
        v = C.c_uint64(0)
        _lib.check(self._lib.hfdtd_adjoint_memory_bytes(self._handle, C.byref(v)))
        return int(v.value) + self.fdtd.device_memory_bytes()

    def read_adjoint_volume(self, name: str) -> np.ndarray:
        """Ẽx..H̃z of the reverse sweep as a volume (for tests).  The library
        stores the E adjoint scaled by ce; it is divided out here."""
        which = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz").index(name)
        out = np.empty(self.shape, dtype=self.fdtd.dtype)
        _lib.check(self._lib.hfdtd_adjoint_read_volume(self._handle, which,
                                                       out.ctypes.data_as(C.c_void_p)))  # fmt: skip
        if which < 3:
            return out.astype(np.float64) / self.fdtd.read_volume("ce_" + "xyz"[which])
        return out

    def close(self) -> None:
        if getattr(self, "_handle", None) is not None:
            self._lib.hfdtd_adjoint_destroy(self._handle)
            self._handle = None
        if getattr(self, "fdtd", None) is not None:
            self.fdtd.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _function():
    import torch

    class ProbeDFT(torch.autograd.Function):
        """ε_design → probe DFTs F (complex128); backward by the adjoint sweep."""

        @staticmethod
        def forward(ctx, eps_design, problem):
            eps = eps_design.detach().to("cpu", torch.float64).numpy()
            F = problem.forward(eps)
            ctx.problem = problem
            ctx.token = problem.token
            ctx.eps = eps
            ctx.like = (eps_design.device, eps_design.dtype)
            return torch.from_numpy(F).to(eps_design.device)

        @staticmethod
        def backward(ctx, grad_F):
            problem = ctx.problem
            if problem.token != ctx.token:  # another forward ran since: redo this one
                problem.forward(ctx.eps)
            g = problem.backward(grad_F.detach().to("cpu", torch.complex128).numpy())
            dev, dt = ctx.like
            return torch.from_numpy(g).to(dev, dt), None

    return ProbeDFT


_FN = None


def probe_dft(problem: ProbeDFTAdjoint, eps_design):
    """Differentiable probe DFTs: ``F = probe_dft(problem, eps)`` with ``eps`` a
    real torch tensor over the design box; ``F`` is complex128 (n_probe, n_freq)."""
    global _FN
    if _FN is None:
        _FN = _function()
    return _FN.apply(eps_design, problem)
