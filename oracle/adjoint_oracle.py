"""Discrete adjoint of the oracle's Yee + CPML step — TEST INFRASTRUCTURE ONLY.

Only ``tests/`` may import this module; it is the checker for the HIP
adjoint (``hipfdtd.adjoint``), never the thing measured or shipped.

The reference has no adjoint: this restates, as plain linear algebra, the
transpose of one forward step of :class:`oracle.fdtd_oracle.OracleFDTD`
(``kernels/yee_update.cl:11-238`` with the step order of
``engine.py:724-737``: H and its ψ from E, sources, E and its ψ from H).
With the state x = (E, H, ψ_H, ψ_E) one lossless step is
``x' = A x + s(ce)`` and

* E half:  ψ_E' = b ψ_E + c D'H,  E' = ca E + ce (ik D'H + S ψ_E')
* H half:  ψ_H' = b ψ_H + c D E,  H' = H − dtm (ik D E + S ψ_H')

with D the forward difference that is zero on the last index and D' the
backward difference that is zero on index 0 (``fdtd_oracle._fwd/_bwd``).
The transposes, taken in reverse order, are

* E half:  w = σ ce Ē',  g = ψ̄_E' + w,  H̄ += D'ᵀ(ik w + c g),  ψ̄_E = b g,  Ē = ca Ē'
* H half:  w = −σ dtm H̄,  g = ψ̄_H' + w,  Ē += Dᵀ(ik w + c g),  ψ̄_H = b g

(σ = ±1 is the sign of the term in the curl, g and ψ̄ live on the stripe of
the derivative's axis), with (Dᵀu)_m = [m≥1] u_{m−1} − [m≤N−2] u_m and
(D'ᵀu)_m = [m≥1] u_m − [m≤N−2] u_{m+1}.

The objective is a function L of probe DFTs F_p = Σ_{n=1..N} E_p^n φ_n,
φ_n = e^{jω n dt} dt.  With g_F the conjugate-Wirtinger gradient that torch
hands back (g_F = 2F for L = |F|²), ∂L/∂E_p^n = Re(conj(g_F) φ_n) is the
adjoint source after step n.  Since ce scales both the curl and the current
sources (E += −ce J), ∂E^n/∂ce = (E^n − E^{n−1})/ce for a lossless step, so
dL/dce = G/ce with G = Σ_n Ē^n (E^n − E^{n−1}); ce = dt/(ε0 εe) turns this
into dL/dεe = −G/εe, and the edge average εe (``fdtd_oracle.edge_average``)
is transposed onto the cells.

How it is pinned: ``tests/test_adjoint_oracle.py`` checks the step with the
dot-product test ⟨A x, y⟩ = ⟨x, Aᵀ y⟩ on random states (a wrong sign, index
or stripe fails it at O(1)), and the whole gradient against torch autograd
of an independent FP64 forward restatement and against central differences.
"""

from __future__ import annotations

import numpy as np

from .fdtd_oracle import MU0, OracleFDTD, _bc

# (field component, derivative axis, sign in the curl, differentiated
# component, psi key) for the E and H updates of fdtd_oracle.update_E/H
E_TERMS = (
    (0, 1, +1, 2, "Ex_y"), (0, 2, -1, 1, "Ex_z"),
    (1, 2, +1, 0, "Ey_z"), (1, 0, -1, 2, "Ey_x"),
    (2, 0, +1, 1, "Ez_x"), (2, 1, -1, 0, "Ez_y"),
)  # fmt: skip
H_TERMS = (
    (0, 1, +1, 2, "Hx_y"), (0, 2, -1, 1, "Hx_z"),
    (1, 2, +1, 0, "Hy_z"), (1, 0, -1, 2, "Hy_x"),
    (2, 0, +1, 1, "Hz_x"), (2, 1, -1, 0, "Hz_y"),
)  # fmt: skip
PSI_KEYS = tuple(t[4] for t in H_TERMS + E_TERMS)


def _sl(axis, s):
    out = [slice(None)] * 3
    out[axis] = s
    return tuple(out)


def fwd_T(u: np.ndarray, axis: int) -> np.ndarray:
    """Transpose of the forward difference: [m≥1] u[m−1] − [m≤N−2] u[m]."""
    out = np.zeros_like(u)
    out[_sl(axis, slice(1, None))] += u[_sl(axis, slice(None, -1))]
    out[_sl(axis, slice(None, -1))] -= u[_sl(axis, slice(None, -1))]
    return out


def bwd_T(u: np.ndarray, axis: int) -> np.ndarray:
    """Transpose of the backward difference: [m≥1] u[m] − [m≤N−2] u[m+1]."""
    out = np.zeros_like(u)
    out[_sl(axis, slice(1, None))] += u[_sl(axis, slice(1, None))]
    out[_sl(axis, slice(None, -1))] -= u[_sl(axis, slice(1, None))]
    return out


def stripe_mask(shape, npml: int, axis: int) -> np.ndarray:
    """Cells of the two face stripes normal to ``axis`` (engine.py:290-318)."""
    n = shape[axis]
    m = np.zeros(n, dtype=bool)
    if npml > 0:
        m[: min(npml, n)] = True
        m[max(n - npml, 0) :] = True
    return _bc(m, axis)


class AdjointState:
    """λ = (Ē, H̄, ψ̄) with ψ̄ kept as volumes that are zero off their stripe."""

    def __init__(self, shape, dtype=np.float64):
        z = lambda: np.zeros(shape, dtype=dtype)  # noqa: E731
        self.E = [z(), z(), z()]
        self.H = [z(), z(), z()]
        self.psi = {k: z() for k in PSI_KEYS}


def psi_volumes(sim: OracleFDTD) -> dict:
    """The oracle's face-stripe ψ as zero-padded volumes."""
    out = {}
    for k, st in sim.psi.items():
        vol = np.zeros(sim.shape, dtype=sim.dtype)
        axis = "xyz".index(k[-1])
        if sim.npml > 0:
            for dom, s in sim._stripes(axis):
                vol[_sl(axis, dom)] = st[_sl(axis, s)]
        out[k] = vol
    return out


def adjoint_step(sim: OracleFDTD, lam: AdjointState) -> None:
    """λ ← Aᵀ λ for one step of ``sim`` (coefficients and profiles only)."""
    dtm = sim.dt / MU0
    shape, n = sim.shape, sim.npml
    masks = [stripe_mask(shape, n, a) for a in range(3)]
    E, H, psi = lam.E, lam.H, lam.psi
    # E half, transposed: every term sees the same Ē'
    for e, a, sg, h, key in E_TERMS:
        b, c, ik = (_bc(sim.prof_e[a][q].astype(np.float64), a) for q in range(3))
        w = sg * sim.ce[e] * E[e]
        g = (psi[key] + w) * masks[a]
        H[h] = H[h] + bwd_T(ik * w + c * g, a)
        psi[key] = b * g
    for e in range(3):
        E[e] = sim.ca[e] * E[e]
    # H half, transposed, with the H̄ just formed
    for h, a, sg, e, key in H_TERMS:
        b, c, ik = (_bc(sim.prof_h[a][q].astype(np.float64), a) for q in range(3))
        w = -sg * dtm * H[h]
        g = (psi[key] + w) * masks[a]
        E[e] = E[e] + fwd_T(ik * w + c * g, a)
        psi[key] = b * g


def edge_average_T(ge, axis: int) -> np.ndarray:
    """Transpose of ``fdtd_oracle.edge_average`` along one axis."""
    out = np.zeros_like(ge)
    out[_sl(axis, slice(None, -1))] += 0.5 * ge[_sl(axis, slice(None, -1))]
    out[_sl(axis, slice(-1, None))] += ge[_sl(axis, slice(-1, None))]
    out[_sl(axis, slice(1, None))] += 0.5 * ge[_sl(axis, slice(None, -1))]
    return out


def source_table(sources, steps: int, dt: float) -> np.ndarray:
    """amps[n, q] = waveform_q((n + ½) dt): step n's current at t + dt/2."""
    t = (np.arange(steps, dtype=np.float64) + 0.5) * dt
    return np.array([[float(src[2](tt)) for src in sources] for tt in t]).reshape(steps, len(sources))


def phases(omegas, n, dt):
    """φ_n = e^{jω n dt} dt for each ω."""
    return np.exp(1j * np.asarray(omegas, dtype=np.float64) * (n * dt)) * dt


def probe_dft_gradient(shape, dl, npml, eps, sources, probes, omegas, steps, grad_F=None,
                       dtype=np.float64, return_state=False):  # fmt: skip
    """Probe DFTs F and dL/dε for every cell, by forward + reverse sweep.

    sources: [(component, (i0,i1,j0,j1,k0,k1), waveform(t))] current sources;
    probes: [(component, i, j, k)]; F has shape (n_probe, n_freq); grad_F is
    the conjugate-Wirtinger gradient of L at F (default 2F, L = Σ|F|²).
    Returns (F, dL/dε as a volume), plus the adjoint state λ⁰ (before step
    1) when ``return_state``.
    """
    sim = OracleFDTD(shape, dl, npml=npml, dtype=dtype)
    sim.set_epsilon(np.asarray(eps, dtype=np.float64))
    amps = source_table(sources, steps, sim.dt)
    for q, (comp, box, _) in enumerate(sources):
        sim.add_source(lambda f, q=q, comp=comp, box=box: f.add_source_J(comp, amps[f.step_num, q], box))
    hist = [[f.astype(np.float64).copy() for f in (sim.Ex, sim.Ey, sim.Ez)]]
    F = np.zeros((len(probes), len(omegas)), dtype=np.complex128)
    for n in range(1, steps + 1):
        sim.step()
        Es = (sim.Ex, sim.Ey, sim.Ez)
        hist.append([f.astype(np.float64).copy() for f in Es])
        ph = phases(omegas, n, sim.dt)
        for p, (c, i, j, k) in enumerate(probes):
            F[p] += float(Es[c][i, j, k]) * ph
    gF = 2.0 * F if grad_F is None else np.asarray(grad_F, dtype=np.complex128)
    lam = AdjointState(sim.shape)
    G = [np.zeros(sim.shape) for _ in range(3)]

    def inject(n):
        ph = phases(omegas, n, sim.dt)
        for p, (c, i, j, k) in enumerate(probes):
            lam.E[c][i, j, k] += float(np.sum((np.conj(gF[p]) * ph).real))

    inject(steps)
    for n in range(steps, 0, -1):
        for c in range(3):
            G[c] += lam.E[c] * (hist[n][c] - hist[n - 1][c])
        adjoint_step(sim, lam)
        if n - 1 >= 1:
            inject(n - 1)
    from .fdtd_oracle import edge_average

    grad = np.zeros(sim.shape)
    for c, ee in enumerate(edge_average(np.asarray(eps, dtype=np.float64))):
        grad += edge_average_T(-G[c] / ee, c)
    if return_state:
        return F, grad, lam
    return F, grad
