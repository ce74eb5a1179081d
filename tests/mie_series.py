"""Mie scattering of a homogeneous sphere — test infrastructure.

Written from the textbook series (Bohren & Huffman, *Absorption and
Scattering of Light by Small Particles*, 1983, §4.4 and §4.8): with the
Riccati–Bessel functions ψ_n(z) = z j_n(z) and ξ_n(z) = z h_n^(1)(z),

    a_n = (m ψ_n(mx) ψ_n'(x) − ψ_n(x) ψ_n'(mx)) / (m ψ_n(mx) ξ_n'(x) − ξ_n(x) ψ_n'(mx))
    b_n = (ψ_n(mx) ψ_n'(x) − m ψ_n(x) ψ_n'(mx)) / (ψ_n(mx) ξ_n'(x) − m ξ_n(x) ψ_n'(mx))

and the angular functions π_n, τ_n of μ = cos θ by their upward
recurrences, S1 = Σ (2n+1)/(n(n+1)) (a_n π_n + b_n τ_n) and
S2 = Σ (2n+1)/(n(n+1)) (a_n τ_n + b_n π_n).  The E-plane intensity is
|S2|².  The series is cut after n = x + 4 x^{1/3} + 2 terms (Wiscombe's
criterion).  The reference keeps its own helper (``tests/analytic_mie.py``)
for the same check (``tests/test_analytic_validation.py:353-411``); this one
is independent and uses scipy's spherical Bessel functions.
"""

from __future__ import annotations

import numpy as np
from scipy.special import spherical_jn, spherical_yn


def n_terms(x: float) -> int:
    return int(np.ceil(x + 4.0 * x ** (1.0 / 3.0) + 2.0))


def _psi(n, z):
    return z * spherical_jn(n, z)


def _dpsi(n, z):
    return spherical_jn(n, z) + z * spherical_jn(n, z, derivative=True)


def _xi(n, z):
    return z * (spherical_jn(n, z) + 1j * spherical_yn(n, z))


def _dxi(n, z):
    h = spherical_jn(n, z) + 1j * spherical_yn(n, z)
    dh = spherical_jn(n, z, derivative=True) + 1j * spherical_yn(n, z, derivative=True)
    return h + z * dh


def coefficients(m, x, nmax=None):
    """(a_n, b_n) for n = 1..nmax; ``m`` relative refractive index, ``x = k a``."""
    nmax = n_terms(x) if nmax is None else int(nmax)
    n = np.arange(1, nmax + 1)
    mx = m * x
    pm, dpm = _psi(n, mx), _dpsi(n, mx)
    px, dpx = _psi(n, x), _dpsi(n, x)
    xx, dxx = _xi(n, x), _dxi(n, x)
    a = (m * pm * dpx - px * dpm) / (m * pm * dxx - xx * dpm)
    b = (pm * dpx - m * px * dpm) / (pm * dxx - m * xx * dpm)
    return a, b


def amplitudes(m, x, theta):
    """S1(θ), S2(θ) for scattering angles ``theta`` [rad]."""
    a, b = coefficients(m, x)
    mu = np.cos(np.asarray(theta, dtype=np.float64))
    s1 = np.zeros(mu.shape, dtype=np.complex128)
    s2 = np.zeros(mu.shape, dtype=np.complex128)
    pi_prev, pi_cur = np.zeros_like(mu), np.ones_like(mu)  # π_0, π_1
    for n in range(1, len(a) + 1):
        tau = n * mu * pi_cur - (n + 1) * pi_prev
        w = (2 * n + 1) / (n * (n + 1))
        s1 += w * (a[n - 1] * pi_cur + b[n - 1] * tau)
        s2 += w * (a[n - 1] * tau + b[n - 1] * pi_cur)
        pi_prev, pi_cur = pi_cur, ((2 * n + 1) * mu * pi_cur - (n + 1) * pi_prev) / n
    return s1, s2


def eplane_intensity(m, x, theta):
    """|S2(θ)|²: scattered intensity in the plane of the incident E field."""
    return np.abs(amplitudes(m, x, theta)[1]) ** 2


def efficiencies(m, x):
    """(Q_ext, Q_sca) of the sphere."""
    a, b = coefficients(m, x)
    n = np.arange(1, len(a) + 1)
    qsca = 2.0 / x**2 * float(np.sum((2 * n + 1) * (np.abs(a) ** 2 + np.abs(b) ** 2)))
    qext = 2.0 / x**2 * float(np.sum((2 * n + 1) * np.real(a + b)))
    return qext, qsca
