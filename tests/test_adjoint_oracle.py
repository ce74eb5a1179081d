"""CPU: the oracle's discrete adjoint and its gradient.

* dot-product test ⟨A x, y⟩ = ⟨x, Aᵀ y⟩ of one step on random states, with
  and without loss (pins every sign, difference transpose and ψ stripe of
  ``oracle/adjoint_oracle.adjoint_step``);
* the torch FP64 forward restatement (``tests/torch_fdtd.py``) reproduces
  the oracle's probe DFTs (FP64, 1e-12);
* the oracle's adjoint gradient equals torch autograd of that restatement
  on every cell (1e-9 peak-normalised) and central differences of the
  oracle objective (δε = 1e-3, 1e-5 relative).
"""

from __future__ import annotations

import numpy as np
import torch

from oracle import adjoint_oracle as ao
from oracle import fdtd_oracle as fo
from torch_fdtd import probe_dft

SHAPE = (12, 13, 14)
NPML = 3
DL = 1e-3
F0 = 3.0e10


def waveform(t):
    return fo.gaussian_sine(t, F0, F0)


SOURCES = [("x", (4, 5, 6, 7, 5, 6), waveform), ("z", (6, 7, 4, 5, 8, 9), lambda t: 0.5 * waveform(t))]
PROBES = [(0, 7, 7, 9), (0, 7, 8, 9), (2, 8, 7, 9)]
OMEGAS = [2 * np.pi * F0, 2 * np.pi * 0.8 * F0]
STEPS = 60


def eps_field(seed=1):
    rng = np.random.default_rng(seed)
    return rng.uniform(1.0, 4.0, SHAPE)


def random_state(sim, rng):
    for name in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz"):
        setattr(sim, name, rng.standard_normal(SHAPE))
    for k, v in sim.psi.items():
        sim.psi[k] = rng.standard_normal(v.shape)


def dot_state(E, H, psi, lam):
    s = sum(float(np.sum(E[c] * lam.E[c])) + float(np.sum(H[c] * lam.H[c])) for c in range(3))
    return s + sum(float(np.sum(psi[k] * lam.psi[k])) for k in ao.PSI_KEYS)


def check_dot_product(lossy):
    rng = np.random.default_rng(7 if lossy else 3)
    sim = fo.OracleFDTD(SHAPE, DL, npml=NPML, dtype=np.float64)
    eps = eps_field()
    sim.set_epsilon(eps, rng.uniform(0.0, 5.0, SHAPE) if lossy else None)
    random_state(sim, rng)
    x_E = [sim.Ex.copy(), sim.Ey.copy(), sim.Ez.copy()]
    x_H = [sim.Hx.copy(), sim.Hy.copy(), sim.Hz.copy()]
    x_psi = ao.psi_volumes(sim)
    sim.update_H()
    sim.update_E()
    y = ao.AdjointState(SHAPE)
    for c in range(3):
        y.E[c] = rng.standard_normal(SHAPE)
        y.H[c] = rng.standard_normal(SHAPE)
    for k in ao.PSI_KEYS:
        y.psi[k] = rng.standard_normal(SHAPE) * ao.stripe_mask(SHAPE, NPML, "xyz".index(k[-1]))
    lhs = dot_state([sim.Ex, sim.Ey, sim.Ez], [sim.Hx, sim.Hy, sim.Hz], ao.psi_volumes(sim), y)
    ao.adjoint_step(sim, y)
    rhs = dot_state(x_E, x_H, x_psi, y)
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs), (lhs, rhs)


def test_adjoint_step_dot_product_lossless():
    check_dot_product(False)


def test_adjoint_step_dot_product_lossy():
    check_dot_product(True)


def test_adjoint_step_dot_product_fails_on_planted_sign_error():
    """A flipped sign in one ψ term breaks the identity at O(1)."""
    orig = ao.E_TERMS
    try:
        ao.E_TERMS = tuple((e, a, -sg if key == "Ey_x" else sg, h, key) for e, a, sg, h, key in orig)
        try:
            check_dot_product(False)
        except AssertionError:
            return
        raise AssertionError("planted sign error went unnoticed")
    finally:
        ao.E_TERMS = orig


def torch_reference(eps):
    t = torch.tensor(eps, dtype=torch.float64, requires_grad=True)
    F = probe_dft(t, SHAPE, DL, NPML, SOURCES, PROBES, OMEGAS, STEPS)
    J = (F.abs() ** 2).sum()
    J.backward()
    return F.detach().numpy(), t.grad.numpy()


def test_torch_forward_matches_oracle():
    eps = eps_field()
    F_ora, _ = ao.probe_dft_gradient(SHAPE, DL, NPML, eps, SOURCES, PROBES, OMEGAS, 40)
    t = torch.tensor(eps, dtype=torch.float64)
    F_t = probe_dft(t, SHAPE, DL, NPML, SOURCES, PROBES, OMEGAS, 40).numpy()
    assert np.max(np.abs(F_t - F_ora)) <= 1e-12 * np.max(np.abs(F_ora))


def test_oracle_gradient_matches_torch_autograd():
    eps = eps_field()
    F_ora, g_ora = ao.probe_dft_gradient(SHAPE, DL, NPML, eps, SOURCES, PROBES, OMEGAS, STEPS)
    F_t, g_t = torch_reference(eps)
    assert np.max(np.abs(F_t - F_ora)) <= 1e-12 * np.max(np.abs(F_ora))
    err = fo.peak_normalised_error(g_ora, g_t)
    print(f"oracle adjoint vs torch autograd: {err:.3e}")
    assert err < 1e-9
    assert np.max(np.abs(g_t)) > 0


def test_oracle_gradient_matches_central_differences():
    eps = eps_field()
    _, g = ao.probe_dft_gradient(SHAPE, DL, NPML, eps, SOURCES, PROBES, OMEGAS, STEPS)
    cells = [tuple(int(v) for v in np.unravel_index(np.argmax(np.abs(g)), SHAPE)), (6, 7, 7)]
    d = 1e-3
    for cell in cells:
        js = []
        for s in (+1, -1):
            e2 = eps.copy()
            e2[cell] += s * d
            F, _ = ao.probe_dft_gradient(SHAPE, DL, NPML, e2, SOURCES, PROBES, OMEGAS, STEPS)
            js.append(float(np.sum(np.abs(F) ** 2)))
        fd = (js[0] - js[1]) / (2 * d)
        assert abs(fd - g[cell]) <= 1e-5 * max(abs(fd), 1e-3 * np.max(np.abs(g))), (cell, fd, g[cell])
