"""FP64 torch restatement of the forward run, for autograd reference gradients.

Test helper (not a test module).  A second, independent statement of the
oracle's lossless Yee + CPML step (``oracle/fdtd_oracle.py``, which follows
``kernels/yee_update.cl:11-238`` and ``engine.py:724-737``): ψ lives on
whole volumes (c = 0 and b = 1 off the PML keep it zero there), every cell
uses the κ-scaled form, materials enter through the differentiable edge
average ``ce = dt/(ε0 εe)``, current sources add ``−ce J`` on their boxes,
and the probe DFTs are ``F_p = Σ_n E_p^n e^{jω n dt} dt``.  torch autograd of
this forward is the reference dL/dε for the hand-written adjoint.
"""

from __future__ import annotations

import math

import numpy as np
import torch

from oracle.fdtd_oracle import C0, EPS0, MU0, cpml_profile


def _diff_fwd(f, a):
    """Forward difference, zero on the last index."""
    d = torch.diff(f, dim=a)
    pad = torch.zeros_like(f.narrow(a, 0, 1))
    return torch.cat([d, pad], dim=a)


def _diff_bwd(f, a):
    """Backward difference, zero on index 0."""
    d = torch.diff(f, dim=a)
    pad = torch.zeros_like(f.narrow(a, 0, 1))
    return torch.cat([pad, d], dim=a)


def _edge(eps, a):
    n = eps.shape[a]
    if n == 1:
        return eps
    avg = 0.5 * (eps.narrow(a, 0, n - 1) + eps.narrow(a, 1, n - 1))
    return torch.cat([avg, eps.narrow(a, n - 1, 1)], dim=a)


def _bcast(v, a, device):
    shape = [1, 1, 1]
    shape[a] = -1
    return torch.as_tensor(np.asarray(v, dtype=np.float64), device=device).reshape(shape)


def probe_dft(eps, shape, dl, npml, sources, probes, omegas, steps):
    """Complex128 tensor (n_probe, n_freq) of probe DFTs, differentiable in ``eps``.

    sources: [(component, (i0,i1,j0,j1,k0,k1), waveform)] with the step-n
    current taken at (n + ½) dt; probes: [(component, i, j, k)].
    """
    dev = eps.device
    dt = 0.99 * dl / (C0 * math.sqrt(3.0))
    dtm = dt / MU0
    prof = {}
    for a, n in enumerate(shape):
        for stag, off in (("e", 0.0), ("h", 0.5)):
            b, c, kap = cpml_profile(n, npml, dl, dt, off, np.float64)
            prof[stag, a] = (_bcast(b, a, dev), _bcast(c, a, dev),
                             _bcast(1.0 / (kap * dl), a, dev))  # fmt: skip
    ce = [(dt / EPS0) / _edge(eps, a) for a in range(3)]
    z = torch.zeros(shape, dtype=torch.float64, device=dev)
    E = [z, z, z]
    H = [z, z, z]
    psi = {}
    boxes = []
    for comp, box, _ in sources:
        m = torch.zeros(shape, dtype=torch.float64, device=dev)
        i0, i1, j0, j1, k0, k1 = box
        m[i0:i1, j0:j1, k0:k1] = 1.0
        boxes.append(("xyz".index(comp) if isinstance(comp, str) else int(comp), m))
    t_src = (np.arange(steps, dtype=np.float64) + 0.5) * dt
    om = torch.as_tensor(np.asarray(omegas, dtype=np.float64), device=dev)
    F = torch.zeros((len(probes), len(omegas)), dtype=torch.complex128, device=dev)
    pidx = [(c, i, j, k) for c, i, j, k in probes]

    def term(key, stag, a, d):
        b, c, ik = prof[stag, a]
        p = b * psi.get(key, z) + c * d
        psi[key] = p
        return ik * d + p

    for n in range(steps):
        Ex, Ey, Ez = E
        Hx = H[0] - dtm * (term("Hx_y", "h", 1, _diff_fwd(Ez, 1)) - term("Hx_z", "h", 2, _diff_fwd(Ey, 2)))
        Hy = H[1] - dtm * (term("Hy_z", "h", 2, _diff_fwd(Ex, 2)) - term("Hy_x", "h", 0, _diff_fwd(Ez, 0)))
        Hz = H[2] - dtm * (term("Hz_x", "h", 0, _diff_fwd(Ey, 0)) - term("Hz_y", "h", 1, _diff_fwd(Ex, 1)))
        H = [Hx, Hy, Hz]
        E = list(E)
        for q, (c, m) in enumerate(boxes):
            amp = float(sources[q][2](t_src[n]))
            E[c] = E[c] + m * (-ce[c] * amp)
        Ex = E[0] + ce[0] * (term("Ex_y", "e", 1, _diff_bwd(Hz, 1)) - term("Ex_z", "e", 2, _diff_bwd(Hy, 2)))
        Ey = E[1] + ce[1] * (term("Ey_z", "e", 2, _diff_bwd(Hx, 2)) - term("Ey_x", "e", 0, _diff_bwd(Hz, 0)))
        Ez = E[2] + ce[2] * (term("Ez_x", "e", 0, _diff_bwd(Hy, 0)) - term("Ez_y", "e", 1, _diff_bwd(Hx, 1)))
        E = [Ex, Ey, Ez]
        ph = torch.exp(1j * om * ((n + 1) * dt)) * dt
        vals = torch.stack([E[c][i, j, k] for c, i, j, k in pidx])
        F = F + vals[:, None].to(torch.complex128) * ph[None, :]
    return F
