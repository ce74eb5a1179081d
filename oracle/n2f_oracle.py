"""CPU oracle for the near-to-far (N2F) monitor — TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s
``cpu_baseline`` leg may import this module.

From-scratch NumPy restatement of the reference's Huygens-box monitor:

* box snapping with Python round-half-even   — ``monitors.py:28-55``
* face packing x0,x1,y0,y1,z0,z1             — ``monitors.py:396-412``
* per-step co-located tangential DFT          — ``kernels/dft_farfield.cl:116-230``
  with the complex128 phase recurrence       — ``monitors.py:456-459, 571-584``
* relative change vs a snapshot               — ``monitors.py:505-569``
* surface integral N, L (trapezoid weights)   — ``kernels/dft_farfield.cl:372-474``
* N, L → far E, H                             — ``kernels/dft_farfield.cl:559-635``
* XZ polar cut and |S| in dB                  — ``monitors.py:737-755, 807-827``

The L-term sign is a parameter: the OpenCL kernel uses ``η0·N⊥ − r̂×L``
(``dft_farfield.cl:617``), the NumPy monitor ``+ r̂×L`` (``monitors.py:171, 378``).
Which one produced the committed far-field fixtures is settled by
``tests/test_oracle_fixtures.py::test_oracle_matches_farfield_fixtures``: both
are evaluated and only the "minus" form matches.

Pinned by (all in ``tests/``):

* the reference's POCL fixtures ``tests/golden/*_farfield`` (19-angle |S| dB
  cut and the E/H vectors at +z and +x), ``test_oracle_fixtures.py``;
* the phase recurrence against direct evaluation and the box snapping and
  packing arithmetic, ``test_oracle_closed_forms.py``;
* what ``l_sign = +1`` changes: with only H sampled on the faces both signs
  agree, with only E sampled they are opposite,
  ``test_oracle_closed_forms.py::test_l_sign_acts_on_the_magnetic_current_term_only``.

The analytic far-field checks (dipole cos²θ cut, endfire null, Mie E-plane)
run on the HIP engine (``tests/test_gpu_analytic.py``).
"""

from __future__ import annotations

import math

import numpy as np

from .fdtd_oracle import C0, ETA0


def snap_box(shape, dl, center, size):
    """Huygens box indices (monitors.py:37-45): Python ``round`` (half-even)."""
    Nx, Ny, Nz = shape
    (cx, cy, cz), (sx, sy, sz) = center, size

    def idx(v):
        return int(round(v / dl))

    return (
        max(0, idx(cx - sx / 2)), min(Nx - 1, idx(cx + sx / 2)),
        max(0, idx(cy - sy / 2)), min(Ny - 1, idx(cy + sy / 2)),
        max(0, idx(cz - sz / 2)), min(Nz - 1, idx(cz + sz / 2)),
    )  # fmt: skip


def face_layout(box):
    """(counts, offsets, n_face) for the packed x0,x1,y0,y1,z0,z1 faces."""
    ix0, ix1, iy0, iy1, iz0, iz1 = box
    nxf, nyf, nzf = ix1 - ix0 + 1, iy1 - iy0 + 1, iz1 - iz0 + 1
    counts = (nyf * nzf, nyf * nzf, nxf * nzf, nxf * nzf, nxf * nyf, nxf * nyf)
    offs = [0]
    for c in counts[:-1]:
        offs.append(offs[-1] + c)
    return counts, tuple(offs), int(sum(counts))


def face_samples(box):
    """Per packed sample: face id and absolute (i, j, k) (dft_farfield.cl:141-164)."""
    ix0, ix1, iy0, iy1, iz0, iz1 = box
    nxf, nyf, nzf = ix1 - ix0 + 1, iy1 - iy0 + 1, iz1 - iz0 + 1
    faces, I, J, K = [], [], [], []
    for f in range(6):
        if f < 2:
            loc = np.arange(nyf * nzf)
            i = np.full(loc.size, ix0 if f == 0 else ix1)
            j = iy0 + loc // nzf
            k = iz0 + loc % nzf
        elif f < 4:
            loc = np.arange(nxf * nzf)
            i = ix0 + loc // nzf
            j = np.full(loc.size, iy0 if f == 2 else iy1)
            k = iz0 + loc % nzf
        else:
            loc = np.arange(nxf * nyf)
            i = ix0 + loc // nyf
            j = iy0 + loc % nyf
            k = np.full(loc.size, iz0 if f == 4 else iz1)
        faces.append(np.full(loc.size, f))
        I.append(i)
        J.append(j)
        K.append(k)
    return (np.concatenate(faces), np.concatenate(I), np.concatenate(J), np.concatenate(K))


def colocate(fields, shape, face, I, J, K, dtype):
    """Tangential E/H at face centres (dft_farfield.cl:171-227), in ``dtype``.

    Returns a dict comp -> (n_face,) array; components that are normal to a
    face stay zero for that face's samples.
    """
    Nx, Ny, Nz = shape
    r = np.dtype(dtype).type
    half, quarter = r(0.5), r(0.25)
    out = {c: np.zeros(face.size, dtype=dtype) for c in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")}
    F = {c: np.asarray(v, dtype=dtype) for c, v in fields.items()}
    i1 = np.minimum(I + 1, Nx - 1)
    j1 = np.minimum(J + 1, Ny - 1)
    k1 = np.minimum(K + 1, Nz - 1)
    ilo = np.maximum(I - 1, 0)
    jlo = np.maximum(J - 1, 0)
    klo = np.maximum(K - 1, 0)

    def avg2(a, b, ok):
        return np.where(ok, half * (a + b), a)

    def avg4(a, b, c, d):
        return quarter * (((a + b) + c) + d)

    m = face <= 1
    i, j, k = I[m], J[m], K[m]
    out["Ey"][m] = avg2(F["Ey"][i, j, k], F["Ey"][i, j, k1[m]], k1[m] != k)
    out["Ez"][m] = avg2(F["Ez"][i, j, k], F["Ez"][i, j1[m], k], j1[m] != j)
    out["Hy"][m] = avg4(F["Hy"][i, j, k], F["Hy"][ilo[m], j, k], F["Hy"][i, j1[m], k],
                        F["Hy"][ilo[m], j1[m], k])  # fmt: skip
    out["Hz"][m] = avg4(F["Hz"][i, j, k], F["Hz"][ilo[m], j, k], F["Hz"][i, j, k1[m]],
                        F["Hz"][ilo[m], j, k1[m]])  # fmt: skip
    m = (face >= 2) & (face <= 3)
    i, j, k = I[m], J[m], K[m]
    out["Ex"][m] = avg2(F["Ex"][i, j, k], F["Ex"][i, j, k1[m]], k1[m] != k)
    out["Ez"][m] = avg2(F["Ez"][i, j, k], F["Ez"][i1[m], j, k], i1[m] != i)
    out["Hx"][m] = avg4(F["Hx"][i, j, k], F["Hx"][i, jlo[m], k], F["Hx"][i1[m], j, k],
                        F["Hx"][i1[m], jlo[m], k])  # fmt: skip
    out["Hz"][m] = avg4(F["Hz"][i, j, k], F["Hz"][i, jlo[m], k], F["Hz"][i, j, k1[m]],
                        F["Hz"][i, jlo[m], k1[m]])  # fmt: skip
    m = face >= 4
    i, j, k = I[m], J[m], K[m]
    out["Ex"][m] = avg2(F["Ex"][i, j, k], F["Ex"][i, j1[m], k], j1[m] != j)
    out["Ey"][m] = avg2(F["Ey"][i, j, k], F["Ey"][i1[m], j, k], i1[m] != i)
    out["Hx"][m] = avg4(F["Hx"][i, j, k], F["Hx"][i, j, klo[m]], F["Hx"][i, j1[m], k],
                        F["Hx"][i, j1[m], klo[m]])  # fmt: skip
    out["Hy"][m] = avg4(F["Hy"][i, j, k], F["Hy"][i, j, klo[m]], F["Hy"][i1[m], j, k],
                        F["Hy"][i1[m], j, klo[m]])  # fmt: skip
    return out


TANGENTIAL = {0: ("Ey", "Ez", "Hy", "Hz"), 1: ("Ex", "Ez", "Hx", "Hz"), 2: ("Ex", "Ey", "Hx", "Hy")}


class OracleNear2Far:
    """Huygens-box DFT monitor + far-field transform, one or more frequencies.

    ``accum`` is the accumulator dtype (np.float32 reproduces the reference's
    default ``accum_fp64=False``; np.float64 is the FP64 oracle).
    """

    def __init__(self, fdtd, center, size, freqs, accum=np.float64):
        self.fdtd = fdtd
        self.freqs = np.atleast_1d(np.asarray(freqs, dtype=np.float64))
        self.omegas = 2.0 * math.pi * self.freqs
        self.dl = fdtd.dl
        self.shape = (fdtd.Nx, fdtd.Ny, fdtd.Nz)
        self.box = snap_box(self.shape, fdtd.dl, center, size)
        self.counts, self.offsets, self.n_face = face_layout(self.box)
        self.face, self.I, self.J, self.K = face_samples(self.box)
        self.accum = np.dtype(accum)
        nf = self.freqs.size
        self.acc = {
            c: np.zeros((nf, self.n_face, 2), dtype=self.accum)
            for c in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")
        }
        self._phase = None
        self._dphase = np.exp(1j * self.omegas * float(fdtd.dt))
        self._hhalf = np.exp(-0.5j * self.omegas * float(fdtd.dt))
        self._snap = None
        fdtd.add_monitor(self)

    def __call__(self, fdtd):
        if self._phase is None:
            self._phase = np.exp(1j * self.omegas * fdtd.t) * fdtd.dt
        else:
            self._phase = self._phase * self._dphase
        ph_e = self._phase
        ph_h = self._phase * self._hhalf
        vals = colocate(fdtd.fields(), self.shape, self.face, self.I, self.J, self.K, fdtd.dtype)
        a = self.accum.type
        for c, v in vals.items():
            ph = ph_e if c[0] == "E" else ph_h
            va = v.astype(self.accum)
            for f in range(self.freqs.size):
                acc = self.acc[c][f]
                acc[:, 0] = acc[:, 0] + va * a(ph[f].real)
                acc[:, 1] = acc[:, 1] + va * a(ph[f].imag)

    def dft(self, comp, f=0):
        """Complex face-packed DFT of one component (complex128)."""
        a = self.acc[comp][f]
        return a[:, 0].astype(np.float64) + 1j * a[:, 1].astype(np.float64)

    def snapshot(self):
        self._snap = {c: v.copy() for c, v in self.acc.items()}

    def relative_change(self, f=0):
        """sqrt(Σ|cur-prev|² / Σ|cur|²); 1.0 without a snapshot (monitors.py:505-569)."""
        if self._snap is None:
            return 1.0
        num = den = 0.0
        for c in self.acc:
            cur = self.acc[c][f].astype(np.float64)
            prev = self._snap[c][f].astype(np.float64)
            num += float(np.sum((cur - prev) ** 2))
            den += float(np.sum(cur**2))
        if den <= 0.0 or not math.isfinite(den) or not math.isfinite(num):
            return 1.0
        return math.sqrt(max(num, 0.0) / den)

    def wavenumber(self, f=0, k_dtype=np.float64) -> float:
        """k = ω/c0, rounded to ``k_dtype`` first.  The reference rounds k to
        the field dtype (``monitors.py:669``: ``self._real(self.omega / C0)``)."""
        return float(np.dtype(k_dtype).type(self.omegas[f] / C0))

    def nl_integrals(self, obs_points, f=0, k=None):
        """Surface integrals N, L (n_obs, 6) complex128, FP64 arithmetic."""
        pts = np.atleast_2d(np.asarray(obs_points, dtype=np.float64))
        dl = self.dl
        k = self.omegas[f] / C0 if k is None else float(k)
        ix0, ix1, iy0, iy1, iz0, iz1 = self.box
        face, I, J, K = self.face, self.I, self.J, self.K
        xp = np.where(face <= 1, I * dl, (I + 0.5) * dl)
        yp = np.where((face >= 2) & (face <= 3), J * dl, (J + 0.5) * dl)
        zp = np.where(face >= 4, K * dl, (K + 0.5) * dl)
        # trapezoid weights (dft_farfield.cl:431-442)
        ea = np.where(face <= 1, (J == iy0) | (J == iy1), (I == ix0) | (I == ix1))
        eb = np.where(face >= 4, (J == iy0) | (J == iy1), (K == iz0) | (K == iz1))
        w = np.where(ea, 0.5, 1.0) * np.where(eb, 0.5, 1.0) * dl * dl
        nf = np.where(face % 2 == 0, -1.0, 1.0)
        D = {c: self.dft(c, f) for c in self.acc}
        # equivalent currents J = n̂×H, M = −n̂×E (dft_farfield.cl:446-473)
        Jv = np.zeros((3, face.size), dtype=np.complex128)
        Mv = np.zeros((3, face.size), dtype=np.complex128)
        x, y, z = face <= 1, (face >= 2) & (face <= 3), face >= 4
        Jv[1][x], Jv[2][x] = (nf * D["Hz"])[x], (-nf * D["Hy"])[x]
        Mv[1][x], Mv[2][x] = (-nf * D["Ez"])[x], (nf * D["Ey"])[x]
        Jv[0][y], Jv[2][y] = (-nf * D["Hz"])[y], (nf * D["Hx"])[y]
        Mv[0][y], Mv[2][y] = (nf * D["Ez"])[y], (-nf * D["Ex"])[y]
        Jv[0][z], Jv[1][z] = (nf * D["Hy"])[z], (-nf * D["Hx"])[z]
        Mv[0][z], Mv[1][z] = (-nf * D["Ey"])[z], (nf * D["Ex"])[z]
        r = np.linalg.norm(pts, axis=1)
        r = np.maximum(r, 1e-30)
        rh = pts / r[:, None]
        phase = k * (rh[:, 0:1] * xp + rh[:, 1:2] * yp + rh[:, 2:3] * zp)
        e = np.exp(1j * phase) * w
        N = e @ Jv.T
        L = e @ Mv.T
        return np.concatenate([N, L], axis=1), rh, r

    def farfields(self, obs_points, f=0, l_sign=-1.0, k_dtype=np.float64, ang_dtype=np.float64):
        """Far E, H (n_obs, 6) complex128 (dft_farfield.cl:559-635).

        ``k_dtype``/``ang_dtype`` reproduce the reference's rounding of k to the
        field dtype and of k·r to the accumulator dtype (``dft_farfield.cl:581-584``);
        the defaults give the FP64 oracle.
        """
        k = self.wavenumber(f, k_dtype)
        NL, rh, r = self.nl_integrals(obs_points, f, k)
        return nl_to_eh(NL, rh, r, k, l_sign, ang_dtype)

    def farfield_polar_xz(self, distance_m=1000.0, n_angles=73, f=0, l_sign=-1.0,
                          k_dtype=np.float64, ang_dtype=np.float64):  # fmt: skip
        R = float(distance_m)
        if R <= 0.0:
            raise ValueError("distance_m must be positive")
        n = max(3, int(n_angles))
        ang = np.linspace(-180.0, 180.0, n)
        rad = np.deg2rad(ang)
        pts = np.column_stack([R * np.sin(rad), np.zeros(n), R * np.cos(rad)])
        eh = self.farfields(pts, f, l_sign, k_dtype, ang_dtype)
        return ang, np.array([poynting_db(e)[0] for e in eh])


def nl_to_eh(NL, rh, r, k, l_sign=-1.0, ang_dtype=np.float64):
    """E = −jk e^{−jkr}/(4πr) (η0 N⊥ + l_sign r̂×L), H = −(r̂×E)/η0.

    k·r is rounded to ``ang_dtype`` before cos/sin (the reference evaluates it
    in its accumulator precision).
    """
    N, L = NL[:, :3], NL[:, 3:]
    t = np.dtype(ang_dtype).type
    ang = (t(k) * np.asarray(r, dtype=ang_dtype)).astype(np.float64)
    pref = (-1j * k / (4.0 * math.pi * r)) * (np.cos(ang) - 1j * np.sin(ang))
    Nt = N - np.sum(rh * N, axis=1, keepdims=True) * rh
    rxL = np.cross(rh, L)
    E = pref[:, None] * (ETA0 * Nt + l_sign * rxL)
    H = -np.cross(rh, E) / ETA0
    return np.concatenate([E, H], axis=1)


def poynting_mag(ff) -> float:
    """|½ E×H*| (monitors.py:807-813)."""
    E, H = ff[0:3], ff[3:6]
    S = 0.5 * np.cross(E, np.conj(H))
    return float(np.sqrt(np.sum(np.abs(S) ** 2)))


def poynting_db(ff):
    """(20 log10 |S|, |S|); −inf for 0 (monitors.py:816-827)."""
    mag = poynting_mag(ff)
    if not math.isfinite(mag) or mag <= 0.0:
        return float("-inf"), 0.0
    return float(20.0 * math.log10(mag)), mag
