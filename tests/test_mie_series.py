"""The Mie helper against what the mathematics fixes: the Rayleigh limit,
energy conservation for a lossless sphere and S1 = S2 in the forward and
backward directions."""

from __future__ import annotations

import numpy as np
import pytest

import mie_series as mie


def test_rayleigh_limit_pattern_and_first_coefficient():
    m, x = 1.5, 0.01
    a, b = mie.coefficients(m, x)
    # a1 = -(2i/3) x^3 (m^2 - 1)/(m^2 + 2) + O(x^5); b1 = O(x^5)
    want = -2j / 3 * x**3 * (m * m - 1) / (m * m + 2)
    assert abs(a[0] - want) / abs(want) < 1e-3
    assert abs(b[0]) < 1e-3 * abs(a[0])
    theta = np.linspace(0.0, np.pi, 61)
    s2 = mie.eplane_intensity(m, 0.05, theta)
    corr = float(np.corrcoef(s2 / s2.max(), np.cos(theta) ** 2)[0, 1])
    assert corr > 0.999, corr


@pytest.mark.parametrize("m,x", [(1.5, 1.2), (2.0, 1.26), (1.33, 5.0)])
def test_lossless_sphere_extinction_equals_scattering(m, x):
    qext, qsca = mie.efficiencies(m, x)
    assert abs(qext - qsca) / qext < 1e-10


def test_forward_and_backward_amplitudes_agree():
    s1, s2 = mie.amplitudes(2.0, 1.26, np.array([0.0, np.pi]))
    assert abs(s1[0] - s2[0]) < 1e-12 * abs(s1[0])
    assert abs(s1[1] + s2[1]) < 1e-12 * abs(s1[1])  # S1(π) = -S2(π)
