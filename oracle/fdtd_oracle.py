"""CPU oracle for the HIPFDTD hot path — TEST INFRASTRUCTURE ONLY.

This module is the checker, never the thing measured or shipped: only
``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it.  The product path (``hipfdtd``) never imports it
and fails loudly when its HIP library is missing.

It is a from-scratch NumPy restatement of the reference's FP32/FP64 Yee +
face-stripe CPML engine, written in the reference OpenCL kernels' operation
order so that the HIP kernels can be compared with it to a few ulp:

* constants                — ``opencl_fdtd_solver/constants.py:11-14``
* CPML 1-D profiles        — ``opencl_fdtd_solver/cpml.py:68-124`` (E offset 0, H offset ½)
* edge-averaged materials  — ``opencl_fdtd_solver/materials.py:28-89``
* H update                 — ``kernels/yee_update.cl:11-40`` (interior) and ``:42-118`` (PML)
* E update                 — ``kernels/yee_update.cl:120-155`` (interior) and ``:157-238`` (PML)
* face ψ stripes           — ``engine.py:290-318``, ``numpy_face_cpml.py:117-123``
* sources                  — ``engine.py:448-538``, ``kernels/sources.cl:7-49``
* step order / time        — ``engine.py:724-744``
* host constant folding    — ``engine.py:138, 264-266, 541-544, 632-634``

How each part is pinned (see ``tests/test_oracle_*.py``):

* the whole Yee+CPML+Jx path by the reference's POCL FP32 mid-plane fixtures
  (``tests/golden/{vacuum_sheet,dielectric_block}``);
* CPML profiles by the closed forms of ``tests/test_unit_cpml.py:24-50``;
* materials by the closed forms of ``tests/test_materials_lossy.py:17-56``;
* Jx injection by ΔE = −dt/(ε0 εe)·J and the rim weights
  (``tests/test_unit_engine.py:183-243``);
* face ψ ≡ volume ψ (an independent full-volume restatement below,
  ``numpy_engine.py:213-284``), σ = 0 ≡ lossless, and plane-wave energy
  bounds with npml = 0;
* lossy (σ > 0) time stepping, for which the reference holds no fixture, by
  closed forms (``tests/test_oracle_lossy.py``): a uniform field shrinks by
  exactly ca per step and follows exp(-σt/ε) to the Crank–Nicolson error;
  one E update from a hand-set H equals ca·E0 + cb·(CPML-weighted curl)
  term by term in interior and PML cells; a plane wave in a lossy half-space
  decays and turns as the discrete scheme's dispersion relation predicts.
  Planted mistakes (ca dropped, ca on the curl term, the loss not halved,
  the lossless ce for cb) fail these checks;
* vacuum dispersion: a plane wave's phase per wavelength follows the Yee
  relation sin(k dl/2) = sin(ω dt/2) dl/(c dt).
"""

from __future__ import annotations

import math

import numpy as np

C0 = 299_792_458.0
MU0 = 4e-7 * math.pi
EPS0 = 1.0 / (MU0 * C0**2)
ETA0 = math.sqrt(MU0 / EPS0)

CPML_M = 3
CPML_KAPPA_MAX = 15.0
CPML_ALPHA_MAX_OVER_ETA0 = 0.05


# ───────────────────────────── CPML profiles ──────────────────────────────


def cpml_sigma_max(dl: float, m: int = CPML_M) -> float:
    """Taflove optimum σ_max = 0.8(m+1)/(η0 Δ) (cpml.py:68-70)."""
    return 0.8 * (m + 1) / (ETA0 * float(dl))


def cpml_profile(n: int, npml: int, dl: float, dt: float, offset: float, dtype=np.float32):
    """1-D CFS-CPML (b, c, κ) of length n (cpml.py:84-124).

    Depth ξ is 1 at the outer wall: lo side ``(npml - i - offset)/npml``,
    hi side ``(i + 1 - offset)/npml`` for the i-th PML cell counted from the
    interface.  σ = σmax ξ^m, κ = 1 + (κmax-1) ξ^m, α = αmax (1 - ξ);
    b = exp(-(σ/κ + α) dt/ε0) is rounded to ``dtype`` before
    c = σ/κ (b - 1)/(σ + κα)/dl uses it.
    """
    b = np.ones(n, dtype=dtype)
    c = np.zeros(n, dtype=dtype)
    kap = np.ones(n, dtype=dtype)
    if npml <= 0 or n <= 0:
        return b, c, kap
    smax = cpml_sigma_max(dl)
    amax = CPML_ALPHA_MAX_OVER_ETA0 / ETA0
    for i in range(npml):
        for lo, idx in ((True, i), (False, n - npml + i)):
            if idx < 0 or idx >= n:
                continue
            xi = (npml - i - offset) / npml if lo else (i + 1 - offset) / npml
            xi = max(xi, 0.0)
            sig = smax * xi**CPML_M
            kp = 1.0 + (CPML_KAPPA_MAX - 1.0) * xi**CPML_M
            al = amax * (1.0 - xi)
            bv = dtype(math.exp(-(sig / kp + al) * dt / EPS0))
            b[idx] = bv
            den = sig + kp * al
            c[idx] = 0.0 if den == 0.0 else sig / kp * (float(bv) - 1.0) / den / dl
            kap[idx] = kp
    return b, c, kap


def inv_kappa_dl(kap: np.ndarray, dl: float, dtype) -> np.ndarray:
    """``(real)(1/(κ·dl))`` evaluated in FP64 from the stored κ (engine.py:264-266)."""
    return (1.0 / (kap.astype(np.float64) * dl)).astype(dtype)


# ───────────────────────────── materials ──────────────────────────────────


def edge_average(cell: np.ndarray):
    """Cell-centred values averaged onto Ex/Ey/Ez edges (materials.py:28-42).

    Ex sits at (i+½, j, k): mean of cells i and i+1; the last plane keeps the
    cell value.  Likewise Ey along j and Ez along k.
    """
    a = np.asarray(cell, dtype=np.float64)
    ex = a.copy()
    ex[:-1] = 0.5 * (a[:-1] + a[1:])
    ey = a.copy()
    ey[:, :-1] = 0.5 * (a[:, :-1] + a[:, 1:])
    ez = a.copy()
    ez[:, :, :-1] = 0.5 * (a[:, :, :-1] + a[:, :, 1:])
    return ex, ey, ez


def edge_ce(eps_r: np.ndarray, dt: float, dtype=np.float32):
    """ce = dt/(ε0 εe) per component (materials.py:45-56)."""
    s = float(dt) / EPS0
    return tuple((s / e).astype(dtype) for e in edge_average(eps_r))


def edge_ca_cb(eps_r: np.ndarray, sigma: np.ndarray, dt: float, dtype=np.float32):
    """Semi-implicit lossy coefficients (materials.py:59-89).

    loss = σe dt/(2 ε0 εe); ca = (1-loss)/(1+loss); cb = dt/(ε0 εe)/(1+loss).
    """
    out_a, out_b = [], []
    for e, s in zip(edge_average(eps_r), edge_average(sigma)):
        loss = s * float(dt) / (2.0 * EPS0 * e)
        den = 1.0 + loss
        out_a.append(((1.0 - loss) / den).astype(dtype))
        out_b.append(((float(dt) / (EPS0 * e)) / den).astype(dtype))
    return tuple(out_a), tuple(out_b)


# ───────────────────────────── the engine ─────────────────────────────────


def _fwd(f: np.ndarray, axis: int) -> np.ndarray:
    """Forward difference, zero on the last index (yee_update.cl:136-141)."""
    d = np.zeros_like(f)
    hi = [slice(None)] * 3
    lo = [slice(None)] * 3
    hi[axis] = slice(1, None)
    lo[axis] = slice(None, -1)
    d[tuple(lo)] = f[tuple(hi)] - f[tuple(lo)]
    return d


def _bwd(f: np.ndarray, axis: int) -> np.ndarray:
    """Backward difference, zero on index 0 (yee_update.cl:256-261)."""
    d = np.zeros_like(f)
    hi = [slice(None)] * 3
    lo = [slice(None)] * 3
    hi[axis] = slice(1, None)
    lo[axis] = slice(None, -1)
    d[tuple(hi)] = f[tuple(hi)] - f[tuple(lo)]
    return d


def _bc(v: np.ndarray, axis: int) -> np.ndarray:
    shape = [1, 1, 1]
    shape[axis] = -1
    return v.reshape(shape)


class OracleFDTD:
    """Face-stripe CPML Yee engine with the reference's public surface.

    ``OpenCLFDTD(shape, dl, npml=20, dtype=np.float32)`` semantics
    (engine.py:99-210): C-order (Nx, Ny, Nz) arrays, fields start at zero,
    ``ce = dt/ε0`` and ``ca = 1`` until :meth:`set_epsilon`.
    """

    def __init__(self, shape, dl, npml=20, dtype=np.float32, accum_fp64=False):
        self.Nx, self.Ny, self.Nz = (int(s) for s in shape)
        self.shape = (self.Nx, self.Ny, self.Nz)
        self.dl = float(dl)
        self.npml = int(npml)
        self.dtype = np.dtype(dtype)
        if self.dtype not in (np.dtype(np.float32), np.dtype(np.float64)):
            raise ValueError("oracle supports float32 or float64")
        self.real = self.dtype.type
        self.accum_fp64 = bool(accum_fp64) or self.dtype == np.float64
        self.accreal = np.float64 if self.accum_fp64 else self.real
        self.t = 0.0
        self.step_num = 0
        self.dt = 0.99 * self.dl / (C0 * math.sqrt(3.0))
        self.size = self.Nx * self.Ny * self.Nz
        z = lambda: np.zeros(self.shape, dtype=self.dtype)  # noqa: E731
        self.Ex, self.Ey, self.Ez = z(), z(), z()
        self.Hx, self.Hy, self.Hz = z(), z(), z()
        ce0 = self.real(self.dt / EPS0)
        self.ce = [np.full(self.shape, ce0, dtype=self.dtype) for _ in range(3)]
        self.ca = [np.ones(self.shape, dtype=self.dtype) for _ in range(3)]
        self._sources = []
        self._monitors = []
        self._build_cpml()

    # plugin API (plugin.py:23-53)
    def add_source(self, src):
        if not callable(src):
            raise TypeError(f"source must be callable, got {type(src)!r}")
        self._sources.append(src)
        return src

    def add_monitor(self, mon):
        if not callable(mon):
            raise TypeError(f"monitor must be callable, got {type(mon)!r}")
        self._monitors.append(mon)
        return mon

    def _build_cpml(self):
        n, dl, dt, dty = self.npml, self.dl, self.dt, self.dtype.type
        self.prof_h = []
        self.prof_e = []
        for ax, N in enumerate(self.shape):
            bh, ch, kh = cpml_profile(N, n, dl, dt, 0.5, dty)
            be, ce, ke = cpml_profile(N, n, dl, dt, 0.0, dty)
            self.prof_h.append((bh, ch, inv_kappa_dl(kh, dl, dty), kh))
            self.prof_e.append((be, ce, inv_kappa_dl(ke, dl, dty), ke))
        Nx, Ny, Nz = self.shape
        if n > 0:
            sx, sy, sz = (2 * n, Ny, Nz), (Nx, 2 * n, Nz), (Nx, Ny, 2 * n)
        else:
            sx = sy = sz = (0, 0, 0)
        self.psi_x_size = int(np.prod(sx)) if n > 0 else 0
        self.psi_y_size = int(np.prod(sy)) if n > 0 else 0
        self.psi_z_size = int(np.prod(sz)) if n > 0 else 0
        zs = lambda s: np.zeros(s, dtype=self.dtype)  # noqa: E731
        # names follow engine.py:305-318: psi_<field>_<axis>
        self.psi = {
            "Hy_x": zs(sx), "Hz_x": zs(sx), "Ey_x": zs(sx), "Ez_x": zs(sx),
            "Hx_y": zs(sy), "Hz_y": zs(sy), "Ex_y": zs(sy), "Ez_y": zs(sy),
            "Hx_z": zs(sz), "Hy_z": zs(sz), "Ex_z": zs(sz), "Ey_z": zs(sz),
        }  # fmt: skip

    def set_epsilon(self, eps_array, sigma_array=None):
        """Upload edge coefficients (engine.py:212-254)."""
        if eps_array.shape != self.shape:
            raise ValueError(
                f"Epsilon shape mismatch: expected {self.shape}, got {tuple(eps_array.shape)}"
            )
        if sigma_array is None:
            self.ce = list(edge_ce(eps_array, self.dt, self.dtype))
            self.ca = [np.ones(self.shape, dtype=self.dtype) for _ in range(3)]
            return
        if sigma_array.shape != self.shape:
            raise ValueError(
                f"Sigma shape mismatch: expected {self.shape}, got {tuple(sigma_array.shape)}"
            )
        ca, cb = edge_ca_cb(eps_array, sigma_array, self.dt, self.dtype)
        self.ca, self.ce = list(ca), list(cb)

    # ── sources (engine.py:448-538, sources.cl:7-49) ──

    def _sheet(self, i0, i1, j0, j1):
        return (
            0 if i0 is None else int(i0),
            self.Nx if i1 is None else int(i1),
            0 if j0 is None else int(j0),
            self.Ny if j1 is None else int(j1),
        )

    def add_source_Ex(self, z_src, amp, i0=None, i1=None, j0=None, j1=None):
        i0, i1, j0, j1 = self._sheet(i0, i1, j0, j1)
        self.Ex[i0:i1, j0:j1, int(z_src)] += self.real(amp)

    def add_source_Jx(self, z_src, Jx, i0=None, i1=None, j0=None, j1=None, *,
                      rim_taper=False, rim_edge=0.8, rim_renorm=True):  # fmt: skip
        i0, i1, j0, j1 = self._sheet(i0, i1, j0, j1)
        jx = float(Jx)
        re = float(rim_edge)
        nx_s, ny_s = max(0, i1 - i0), max(0, j1 - j0)
        if rim_taper and rim_renorm and nx_s >= 2 and ny_s >= 2:
            ni, nj = nx_s - 2, ny_s - 2
            wsum = ni * nj + re * (2 * ni + 2 * nj) + (re * re) * 4
            jx *= (nx_s * ny_s) / wsum
        r = self.real
        z = int(z_src)
        ii = np.arange(max(i0, 0), min(i1, self.Nx))
        jj = np.arange(max(j0, 0), min(j1, self.Ny))
        if ii.size == 0 or jj.size == 0:
            return
        w = np.ones((ii.size, jj.size), dtype=self.dtype)
        if rim_taper:
            edge_i = (ii == i0) | (ii == i1 - 1)
            edge_j = (jj == j0) | (jj == j1 - 1)
            w[edge_i, :] = (w[edge_i, :] * r(re)).astype(self.dtype)
            w[:, edge_j] = (w[:, edge_j] * r(re)).astype(self.dtype)
        sl = (slice(ii[0], ii[-1] + 1), slice(jj[0], jj[-1] + 1), z)
        term = ((-self.ce[0][sl]) * r(jx)) * w
        self.Ex[sl] = self.Ex[sl] + term.astype(self.dtype)

    def add_source_J(self, component, J, box, *, rim_taper=0, rim_edge=0.8):
        """Current density on one E component over a half-open box.

        Generalises add_source_Jx (sources.cl:26-49) to any component and box:
        ``E_c += ((-ce_c) * J) * w`` with w = rim_edge per tapered axis whose
        edge the cell lies on (bit 0: i, bit 1: j, bit 2: k).  Pinned by the
        closed form ΔE = -dt/(ε0 εe) J (tests/test_oracle_closed_forms.py).
        """
        c = "xyz".index(component) if isinstance(component, str) else int(component)
        i0, i1, j0, j1, k0, k1 = (int(v) for v in box)
        r = self.real
        F = (self.Ex, self.Ey, self.Ez)[c]
        ce = self.ce[c]
        for i in range(max(i0, 0), min(i1, self.Nx)):
            for j in range(max(j0, 0), min(j1, self.Ny)):
                for k in range(max(k0, 0), min(k1, self.Nz)):
                    w = r(1.0)
                    if (rim_taper & 1) and i in (i0, i1 - 1):
                        w = r(w * r(rim_edge))
                    if (rim_taper & 2) and j in (j0, j1 - 1):
                        w = r(w * r(rim_edge))
                    if (rim_taper & 4) and k in (k0, k1 - 1):
                        w = r(w * r(rim_edge))
                    F[i, j, k] = F[i, j, k] + ((-ce[i, j, k]) * r(J)) * w

    # ── updates ──

    def _interior_mask(self):
        n = self.npml
        if n <= 0:
            return None
        Nx, Ny, Nz = self.shape
        if Nx - 2 * n <= 0 or Ny - 2 * n <= 0 or Nz - 2 * n <= 0:
            return None
        return (slice(n, Nx - n), slice(n, Ny - n), slice(n, Nz - n))

    def _stripes(self, axis):
        """(domain slice, stripe slice) for the lo and hi PML faces of one axis."""
        n, N = self.npml, self.shape[axis]
        return ((slice(0, n), slice(0, n)), (slice(N - n, N), slice(n, 2 * n)))

    def _psi_update(self, key, axis, prof, d):
        """ψ ← b ψ + c d on the two stripes of ``axis``; returns ψ as a volume."""
        vol = np.zeros(self.shape, dtype=self.dtype)
        if self.npml <= 0:
            return vol
        b, c = prof[0], prof[1]
        psi = self.psi[key]
        for dom, st in self._stripes(axis):
            sel_d = [slice(None)] * 3
            sel_s = [slice(None)] * 3
            sel_d[axis] = dom
            sel_s[axis] = st
            bb = _bc(b[dom], axis)
            cc = _bc(c[dom], axis)
            new = bb * psi[tuple(sel_s)] + cc * d[tuple(sel_d)]
            psi[tuple(sel_s)] = new
            vol[tuple(sel_d)] = new
        return vol

    def update_H(self):
        r = self.real
        dtm = r(self.dt / MU0)
        dtm_dl = r(self.dt / (MU0 * self.dl))
        Ex, Ey, Ez = self.Ex, self.Ey, self.Ez
        dEz_dy, dEy_dz = _fwd(Ez, 1), _fwd(Ey, 2)
        dEx_dz, dEz_dx = _fwd(Ex, 2), _fwd(Ez, 0)
        dEy_dx, dEx_dy = _fwd(Ey, 0), _fwd(Ex, 1)
        ph = self.prof_h
        ikx, iky, ikz = _bc(ph[0][2], 0), _bc(ph[1][2], 1), _bc(ph[2][2], 2)
        p_hx_y = self._psi_update("Hx_y", 1, ph[1], dEz_dy)
        p_hx_z = self._psi_update("Hx_z", 2, ph[2], dEy_dz)
        p_hy_x = self._psi_update("Hy_x", 0, ph[0], dEz_dx)
        p_hy_z = self._psi_update("Hy_z", 2, ph[2], dEx_dz)
        p_hz_x = self._psi_update("Hz_x", 0, ph[0], dEy_dx)
        p_hz_y = self._psi_update("Hz_y", 1, ph[1], dEx_dy)
        # PML form (yee_update.cl:115-117) everywhere, interior overwritten below
        cx = dEz_dy * iky + p_hx_y - dEy_dz * ikz - p_hx_z
        cy = dEx_dz * ikz + p_hy_z - dEz_dx * ikx - p_hy_x
        cz = dEy_dx * ikx + p_hz_x - dEx_dy * iky - p_hz_y
        nHx = self.Hx - dtm * cx
        nHy = self.Hy - dtm * cy
        nHz = self.Hz - dtm * cz
        m = self._interior_mask()
        if m is not None:  # interior form (yee_update.cl:37-39)
            nHx[m] = self.Hx[m] - dtm_dl * (dEz_dy[m] - dEy_dz[m])
            nHy[m] = self.Hy[m] - dtm_dl * (dEx_dz[m] - dEz_dx[m])
            nHz[m] = self.Hz[m] - dtm_dl * (dEy_dx[m] - dEx_dy[m])
        self.Hx, self.Hy, self.Hz = nHx, nHy, nHz

    def update_E(self):
        r = self.real
        inv_dl = r(1.0 / self.dl)
        Hx, Hy, Hz = self.Hx, self.Hy, self.Hz
        dHz_dy, dHy_dz = _bwd(Hz, 1), _bwd(Hy, 2)
        dHx_dz, dHz_dx = _bwd(Hx, 2), _bwd(Hz, 0)
        dHy_dx, dHx_dy = _bwd(Hy, 0), _bwd(Hx, 1)
        pe = self.prof_e
        ikx, iky, ikz = _bc(pe[0][2], 0), _bc(pe[1][2], 1), _bc(pe[2][2], 2)
        p_ex_y = self._psi_update("Ex_y", 1, pe[1], dHz_dy)
        p_ex_z = self._psi_update("Ex_z", 2, pe[2], dHy_dz)
        p_ey_x = self._psi_update("Ey_x", 0, pe[0], dHz_dx)
        p_ey_z = self._psi_update("Ey_z", 2, pe[2], dHx_dz)
        p_ez_x = self._psi_update("Ez_x", 0, pe[0], dHy_dx)
        p_ez_y = self._psi_update("Ez_y", 1, pe[1], dHx_dy)
        (cax, cay, caz), (cex, cey, cez) = self.ca, self.ce
        nEx = cax * self.Ex + cex * (dHz_dy * iky + p_ex_y - dHy_dz * ikz - p_ex_z)
        nEy = cay * self.Ey + cey * (dHx_dz * ikz + p_ey_z - dHz_dx * ikx - p_ey_x)
        nEz = caz * self.Ez + cez * (dHy_dx * ikx + p_ez_x - dHx_dy * iky - p_ez_y)
        m = self._interior_mask()
        if m is not None:  # interior form (yee_update.cl:213-215)
            nEx[m] = cax[m] * self.Ex[m] + cex[m] * inv_dl * (dHz_dy[m] - dHy_dz[m])
            nEy[m] = cay[m] * self.Ey[m] + cey[m] * inv_dl * (dHx_dz[m] - dHz_dx[m])
            nEz[m] = caz[m] * self.Ez[m] + cez[m] * inv_dl * (dHy_dx[m] - dHx_dy[m])
        self.Ex, self.Ey, self.Ez = nEx, nEy, nEz

    def step(self):
        """H → sources at t+dt/2 → E → t += dt → monitors (engine.py:724-737)."""
        self.update_H()
        t_int = self.t
        self.t = t_int + 0.5 * self.dt
        for src in self._sources:
            src(self)
        self.t = t_int
        self.update_E()
        self.t += self.dt
        self.step_num += 1
        for mon in self._monitors:
            mon(self)

    def run(self, n_steps):
        for _ in range(int(n_steps)):
            self.step()

    def fields(self):
        return {"Ex": self.Ex, "Ey": self.Ey, "Ez": self.Ez,
                "Hx": self.Hx, "Hy": self.Hy, "Hz": self.Hz}  # fmt: skip


class OracleVolumeFDTD(OracleFDTD):
    """Independent full-volume-ψ restatement (numpy_engine.py:213-284).

    Here ψ lives on the whole volume and every cell uses the κ-scaled curl
    divided by (κ·dl) in dtype arithmetic.  Agreement with the face-stripe
    engine to round-off pins the stripe indexing (a wrong stripe offset or a
    missed face changes fields at O(1) relative level).
    """

    def _build_cpml(self):
        super()._build_cpml()
        self.vpsi = {k: np.zeros(self.shape, dtype=self.dtype) for k in self.psi}

    def update_H(self):
        dtm = self.dt / MU0
        dl = self.dl
        Ex, Ey, Ez = self.Ex, self.Ey, self.Ez
        d = {"Ez_y": _fwd(Ez, 1), "Ey_z": _fwd(Ey, 2), "Ex_z": _fwd(Ex, 2),
             "Ez_x": _fwd(Ez, 0), "Ey_x": _fwd(Ey, 0), "Ex_y": _fwd(Ex, 1)}  # fmt: skip
        P = self.prof_h
        b = [_bc(P[a][0], a) for a in range(3)]
        c = [_bc(P[a][1], a) for a in range(3)]
        k = [_bc(P[a][3], a) for a in range(3)]
        v = self.vpsi
        v["Hx_y"] = b[1] * v["Hx_y"] + c[1] * d["Ez_y"]
        v["Hx_z"] = b[2] * v["Hx_z"] + c[2] * d["Ey_z"]
        v["Hy_x"] = b[0] * v["Hy_x"] + c[0] * d["Ez_x"]
        v["Hy_z"] = b[2] * v["Hy_z"] + c[2] * d["Ex_z"]
        v["Hz_x"] = b[0] * v["Hz_x"] + c[0] * d["Ey_x"]
        v["Hz_y"] = b[1] * v["Hz_y"] + c[1] * d["Ex_y"]
        self.Hx = (self.Hx - dtm * (d["Ez_y"] / (k[1] * dl) + v["Hx_y"]
                                    - d["Ey_z"] / (k[2] * dl) - v["Hx_z"])).astype(self.dtype)
        self.Hy = (self.Hy - dtm * (d["Ex_z"] / (k[2] * dl) + v["Hy_z"]
                                    - d["Ez_x"] / (k[0] * dl) - v["Hy_x"])).astype(self.dtype)
        self.Hz = (self.Hz - dtm * (d["Ey_x"] / (k[0] * dl) + v["Hz_x"]
                                    - d["Ex_y"] / (k[1] * dl) - v["Hz_y"])).astype(self.dtype)

    def update_E(self):
        dl = self.dl
        Hx, Hy, Hz = self.Hx, self.Hy, self.Hz
        d = {"Hz_y": _bwd(Hz, 1), "Hy_z": _bwd(Hy, 2), "Hx_z": _bwd(Hx, 2),
             "Hz_x": _bwd(Hz, 0), "Hy_x": _bwd(Hy, 0), "Hx_y": _bwd(Hx, 1)}  # fmt: skip
        P = self.prof_e
        b = [_bc(P[a][0], a) for a in range(3)]
        c = [_bc(P[a][1], a) for a in range(3)]
        k = [_bc(P[a][3], a) for a in range(3)]
        v = self.vpsi
        v["Ex_y"] = b[1] * v["Ex_y"] + c[1] * d["Hz_y"]
        v["Ex_z"] = b[2] * v["Ex_z"] + c[2] * d["Hy_z"]
        v["Ey_x"] = b[0] * v["Ey_x"] + c[0] * d["Hz_x"]
        v["Ey_z"] = b[2] * v["Ey_z"] + c[2] * d["Hx_z"]
        v["Ez_x"] = b[0] * v["Ez_x"] + c[0] * d["Hy_x"]
        v["Ez_y"] = b[1] * v["Ez_y"] + c[1] * d["Hx_y"]
        (cax, cay, caz), (cex, cey, cez) = self.ca, self.ce
        self.Ex = (cax * self.Ex + cex * (d["Hz_y"] / (k[1] * dl) + v["Ex_y"]
                                          - d["Hy_z"] / (k[2] * dl) - v["Ex_z"])).astype(self.dtype)
        self.Ey = (cay * self.Ey + cey * (d["Hx_z"] / (k[2] * dl) + v["Ey_z"]
                                          - d["Hz_x"] / (k[0] * dl) - v["Ey_x"])).astype(self.dtype)
        self.Ez = (caz * self.Ez + cez * (d["Hy_x"] / (k[0] * dl) + v["Ez_x"]
                                          - d["Hx_y"] / (k[1] * dl) - v["Ez_y"])).astype(self.dtype)


def gaussian_sine(t: float, freq: float, fwidth: float) -> float:
    """Gaussian-modulated sinusoid used by the reference harness
    (tests/meep_validation/harness.py:70-73): t0 = 5/(π fw), σ = 1/(π fw)."""
    t0 = 5.0 / (math.pi * fwidth)
    sig = 1.0 / (math.pi * fwidth)
    return float(math.exp(-0.5 * ((t - t0) / sig) ** 2) * math.sin(2 * math.pi * freq * t))


def peak_normalised_error(a, b) -> float:
    """max |a - b| / max |b| (the parity figure used throughout the tests)."""
    a = np.asarray(a)
    b = np.asarray(b)
    den = float(np.max(np.abs(b)))
    if den == 0.0:
        return float(np.max(np.abs(a)))
    return float(np.max(np.abs(a - b))) / den
