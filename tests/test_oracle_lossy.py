"""Pins of the oracle's lossy (σ > 0, ca ≠ 1) time step.

The reference's lossy E update is ``Ex = ca·Ex + ce·curl`` in both the
interior and the PML kernels (``kernels/yee_update.cl:152-154, 235-237``),
with the semi-implicit coefficients of ``materials.py:59-89``.  The
reference itself checks it only by invariants (``tests/test_materials_lossy.py:80-123``);
these tests pin the oracle's restatement by what the mathematics fixes:

* a uniform field in a uniform lossy medium has zero curl, so each step
  multiplies it by ca exactly, and ca^n follows exp(-σ t/ε) to the
  Crank–Nicolson truncation error, -(2/3)·n·L³ with L = σ dt/(2 ε0 εr);
* one E update from a hand-set H whose differences are known in closed
  form gives ca·E0 plus the CPML-weighted curl, term by term, in interior
  and PML cells;
* a plane wave in a lossy slab decays and turns in phase as the dispersion
  relation of the discrete scheme predicts,
  sin²(k dl/2) = μ dl² sin(ω dt/2)(e^{jω dt/2} − ca e^{−jω dt/2}) / (2j cb dt),
  which tends to the continuous medium's k = ω √(με(1 − jσ/(ωε))).

Each planted mistake below (ca dropped, ca applied to the curl term, the
loss not halved, the lossless ce used in place of cb) must fail at least
one of these checks by far more than the tolerance.
"""

from __future__ import annotations

import math

import numpy as np
import pytest

from oracle import fdtd_oracle as fo

# ───────────────────────── planted mistakes ───────────────────────────────


class _NoCa(fo.OracleFDTD):
    """Planted: E update without the ca·E term (E = E + cb·curl)."""

    def update_E(self):
        ca = self.ca
        self.ca = [np.ones_like(a) for a in ca]
        try:
            super().update_E()
        finally:
            self.ca = ca


class _CaOnCurl(fo.OracleFDTD):
    """Planted: ca multiplies the curl term too (E = ca·E + ca·cb·curl)."""

    def update_E(self):
        ce = self.ce
        self.ce = [(a * b).astype(self.dtype) for a, b in zip(self.ca, ce)]
        try:
            super().update_E()
        finally:
            self.ce = ce


class _LossNotHalved(fo.OracleFDTD):
    """Planted: loss = σ dt/(ε0 εr) instead of σ dt/(2 ε0 εr)."""

    def set_epsilon(self, eps_array, sigma_array=None):
        super().set_epsilon(eps_array, None if sigma_array is None else 2.0 * sigma_array)


class _CeForCb(fo.OracleFDTD):
    """Planted: the lossless ce = dt/(ε0 εe) multiplies the curl instead of cb."""

    def set_epsilon(self, eps_array, sigma_array=None):
        super().set_epsilon(eps_array, sigma_array)
        self.ce = list(fo.edge_ce(eps_array, self.dt, self.dtype))


PLANTED = [_NoCa, _CaOnCurl, _LossNotHalved, _CeForCb]


# ───────────────────────── 1. uniform decay ───────────────────────────────


def uniform_decay_error(cls, sigma=0.5, eps_r=2.0, steps=40):
    """(max relative deviation from ca^n, deviation of ca^n from exp(-σt/ε)
    relative to the Crank–Nicolson prediction) after ``steps`` steps of a
    uniform unit field (all three E components, interior and PML cells)."""
    shape, npml = (12, 12, 12), 3
    sim = cls(shape, 1e-3, npml=npml, dtype=np.float64)
    sim.set_epsilon(np.full(shape, eps_r), np.full(shape, sigma))
    for name in ("Ex", "Ey", "Ez"):
        setattr(sim, name, np.ones(shape))
    sim.run(steps)
    L = sigma * sim.dt / (2.0 * fo.EPS0 * eps_r)
    ca = (1.0 - L) / (1.0 + L)
    want = ca**steps
    dev = max(float(np.max(np.abs(getattr(sim, n) - want))) for n in ("Ex", "Ey", "Ez")) / want
    h = max(float(np.max(np.abs(getattr(sim, n)))) for n in ("Hx", "Hy", "Hz"))
    assert h == 0.0  # zero curl: H never moves
    # ca^n / exp(-σ n dt/ε) - 1 = exp(-2n(L³/3 + L⁵/5 + ...)) - 1 ≈ -(2/3) n L³
    got = float(sim.Ex[0, 0, 0])
    cn = got / math.exp(-sigma * steps * sim.dt / (fo.EPS0 * eps_r)) - 1.0
    pred = -(2.0 / 3.0) * steps * L**3
    return dev, abs(cn - pred) / abs(pred)


def test_lossy_uniform_field_decays_by_ca_each_step():
    dev, cn = uniform_decay_error(fo.OracleFDTD)
    assert dev < 1e-12, dev
    assert cn < 0.05, cn  # the Crank–Nicolson L³ term, observed 1.7e-4


@pytest.mark.parametrize("cls", [_NoCa, _LossNotHalved])
def test_planted_decay_mistakes_fail(cls):
    _, cn = uniform_decay_error(cls)
    assert cn > 10.0, (cls.__name__, cn)


# ───────────────────────── 2. one step, term by term ──────────────────────


def one_step_error(cls):
    """Peak-normalised error of one E update from hand-set H against the
    closed form ca·E0 + cb·Σ ±d_a·(ik_a + c_a), FP64, every cell."""
    shape, npml, dl = (14, 15, 16), 3, 1e-3
    sim = cls(shape, dl, npml=npml, dtype=np.float64)
    rng = np.random.default_rng(12)
    eps_r = rng.uniform(1.0, 4.0, shape)
    sig = rng.uniform(0.5, 3.0, shape)
    sim.set_epsilon(eps_r, sig)
    e0 = (0.7, -0.4, 1.1)
    for name, v in zip(("Ex", "Ey", "Ez"), e0):
        setattr(sim, name, np.full(shape, v))
    # H linear in the indices: every backward difference is a known constant
    a, a2, q, q2, p, p2 = 3e-3, -2e-3, 1.5e-3, 4e-3, -2.5e-3, 1e-3
    I, J, K = np.indices(shape, dtype=np.float64)
    sim.Hz = a * J + a2 * I
    sim.Hy = q * K + q2 * I
    sim.Hx = p * K + p2 * J
    sim.update_E()

    # closed form; the coefficients come from their own closed forms
    dt = sim.dt
    ex_e, ey_e, ez_e = fo.edge_average(eps_r)
    sx_e, sy_e, sz_e = fo.edge_average(sig)

    def coeffs(ee, se):
        L = se * dt / (2.0 * fo.EPS0 * ee)
        return (1.0 - L) / (1.0 + L), dt / (fo.EPS0 * ee) / (1.0 + L)

    def w(axis, idx):
        # CPML weight of a difference along ``axis``: ik + c (ψ starts at 0)
        N = shape[axis]
        b, c, kap = fo.cpml_profile(N, npml, dl, dt, 0.0, np.float64)
        return (1.0 / (kap * dl) + c)[idx]

    step = lambda X: (X >= 1).astype(np.float64)  # noqa: E731  (zero difference at index 0)
    wi, wj, wk = w(0, I.astype(int)), w(1, J.astype(int)), w(2, K.astype(int))
    want = []
    for (ee, se), v, curl in (
        ((ex_e, sx_e), e0[0], a * step(J) * wj - q * step(K) * wk),
        ((ey_e, sy_e), e0[1], p * step(K) * wk - a2 * step(I) * wi),
        ((ez_e, sz_e), e0[2], q2 * step(I) * wi - p2 * step(J) * wj),
    ):
        ca, cb = coeffs(ee, se)
        want.append(ca * v + cb * curl)
    got = (sim.Ex, sim.Ey, sim.Ez)
    peak = max(float(np.max(np.abs(x))) for x in want)
    return max(float(np.max(np.abs(g - x))) for g, x in zip(got, want)) / peak


def test_lossy_update_E_term_by_term():
    err = one_step_error(fo.OracleFDTD)
    assert err < 1e-12, err


@pytest.mark.parametrize("cls", PLANTED)
def test_planted_one_step_mistakes_fail(cls):
    err = one_step_error(cls)
    assert err > 1e-3, (cls.__name__, err)


# ───────────────────────── 3. plane wave in a lossy slab ──────────────────

SLAB = dict(sigma=3.0, eps_r=2.0, npml=2, nz=200, z_src=30, z_slab=70, steps=1600, lam=40)


def plane_wave(cls, sigma_z, eps_r, nz, steps, dtype=np.float64, stride=1, npml=2, z_src=30,
               lam=40):  # fmt: skip
    """A plane wave along z (Jx on whole xy sheets, so the field is exactly
    uniform in x and y) of wavelength ``lam`` cells in vacuum, in a medium
    of relative permittivity ``eps_r`` and conductivity ``sigma_z`` (one
    value per z plane).  Returns (complex amplitude of Ex along z at the
    source frequency, fitted over the last 400 steps sampled every
    ``stride`` steps, engine, ω)."""
    shape = (2 * npml + 2, 2 * npml + 2, nz)
    sim = cls(shape, 1e-3, npml=npml, dtype=dtype)
    sig = np.broadcast_to(np.asarray(sigma_z, dtype=dtype), shape).copy()
    sim.set_epsilon(np.full(shape, eps_r, dtype), sig)
    f = fo.C0 / (lam * 1e-3)
    w = 2.0 * math.pi * f
    tau = 3.0 / f
    sim.add_source(lambda s: s.add_source_Jx(
        z_src, math.sin(w * s.t) * (1.0 - math.exp(-((s.t / tau) ** 2))), 0, shape[0], 0,
        shape[1]))  # fmt: skip
    rec, tt = [], []
    for s in range(0, steps, stride):
        sim.run(stride)
        if s >= steps - 400:
            rec.append(np.asarray(sim.Ex)[npml, npml, :].astype(np.float64))
            tt.append(sim.t)
    rec, tt = np.array(rec), np.array(tt)
    M = np.column_stack([np.cos(w * tt), np.sin(w * tt)])
    coef = np.linalg.lstsq(M, rec, rcond=None)[0]
    return coef[0] - 1j * coef[1], sim, w  # Ex = Re(amp · e^{jωt})


def plane_wave_in_slab(cls, dtype=np.float64, stride=1, **kw):
    """The plane wave entering a lossy half-space z >= z_slab (``SLAB``)."""
    c = dict(SLAB, **kw)
    sig = np.where(np.arange(c["nz"]) >= c["z_slab"], c["sigma"], 0.0)
    return plane_wave(cls, sig, c["eps_r"], c["nz"], c["steps"], dtype, stride, c["npml"],
                      c["z_src"], c["lam"])  # fmt: skip


def vacuum_phase_error(cls, dtype=np.float64, stride=1):
    """Relative error of the phase a vacuum plane wave turns over one
    wavelength (40 cells, so a reflected wave's ripple cancels) against the
    Yee relation sin(k dl/2) = sin(ω dt/2) dl/(c dt).  The far end absorbs
    through a conductivity graded quadratically from 0 over 160 cells."""
    nz = 300
    z = np.arange(nz)
    sig = np.clip((z - 140) / 160.0, 0.0, 1.0) ** 2
    amp, sim, w = plane_wave(cls, sig, 1.0, nz, 1800, dtype, stride)
    k = 2.0 / sim.dl * math.asin(math.sin(0.5 * w * sim.dt) * sim.dl / (fo.C0 * sim.dt))
    want = k * sim.dl * 40
    worst = 0.0
    for z0 in (60, 67, 73):
        d = float(np.angle(amp[z0 + 40] / amp[z0])) + want
        d = (d + math.pi) % (2.0 * math.pi) - math.pi
        worst = max(worst, abs(d) / want, abs(abs(amp[z0 + 40]) / abs(amp[z0]) - 1.0))
    return worst


def discrete_k(w, dt, dl, sigma, eps_r):
    """Complex wavenumber of the semi-implicit lossy Yee scheme (1-D)."""
    L = sigma * dt / (2.0 * fo.EPS0 * eps_r)
    ca = (1.0 - L) / (1.0 + L)
    cb = dt / (fo.EPS0 * eps_r) / (1.0 + L)
    S = (math.sin(w * dt / 2) * (np.exp(0.5j * w * dt) - ca * np.exp(-0.5j * w * dt))
         * fo.MU0 * dl**2 / (2j * cb * dt))  # fmt: skip
    k = 2.0 / dl * np.arcsin(np.sqrt(S + 0j))
    return complex(k.real, -abs(k.imag))  # e^{-jkz}: decaying towards +z


def slab_errors(amp, sim, w, k):
    """Relative errors of the measured per-cell decay and phase over 20
    cells inside the slab against wavenumber ``k``."""
    z = SLAB["z_slab"] + 5
    decay = math.log(abs(amp[z + 20]) / abs(amp[z])) / 20
    turn = float(np.angle(amp[z + 5] / amp[z])) / 5  # 5 cells: no phase wrap
    want_decay, want_turn = k.imag * sim.dl, -k.real * sim.dl
    return abs(decay - want_decay) / abs(want_decay), abs(turn - want_turn) / abs(want_turn)


def test_lossy_plane_wave_matches_discrete_dispersion():
    amp, sim, w = plane_wave_in_slab(fo.OracleFDTD)
    kd = discrete_k(w, sim.dt, sim.dl, SLAB["sigma"], SLAB["eps_r"])
    ed, et = slab_errors(amp, sim, w, kd)
    assert ed < 1e-4 and et < 1e-4, (ed, et)  # observed 7.5e-6 and 5.9e-7
    # and the discrete scheme tends to the continuous lossy medium
    eps = fo.EPS0 * SLAB["eps_r"]
    kc = w * np.sqrt(fo.MU0 * eps * (1 - 1j * SLAB["sigma"] / (w * eps)))
    ed, et = slab_errors(amp, sim, w, complex(kc.real, -abs(kc.imag)))
    assert ed < 0.03 and et < 0.03, (ed, et)


@pytest.mark.parametrize("cls", PLANTED)
def test_planted_slab_mistakes_fail(cls):
    amp, sim, w = plane_wave_in_slab(cls)
    kd = discrete_k(w, sim.dt, sim.dl, SLAB["sigma"], SLAB["eps_r"])
    ed, et = slab_errors(amp, sim, w, kd)
    assert max(ed, et) > 0.03, (cls.__name__, ed, et)


def test_vacuum_plane_wave_follows_yee_dispersion():
    """σ = 0: the reference computes this relation (test_analytic_validation.py:130-178);
    here the oracle's wave is measured against it."""
    err = vacuum_phase_error(fo.OracleFDTD)
    assert err < 1e-5, err  # observed 5.5e-7
