"""GPU: adjoint ε-gradients of probe-DFT objectives (``hipfdtd.adjoint``).

Bars (peak-normalised over the design box unless stated):

* FP64 engine vs the oracle's discrete adjoint (``oracle/adjoint_oracle``):
  probe DFTs 1e-12, dL/dε 1e-9, final adjoint fields 1e-9;
* FP32 engine vs the same: dL/dε 1e-4;
* 40³ grid, εr ~ U[1, 12] on a 12×12×8 design box: FP64 engine vs torch FP64
  autograd of ``tests/torch_fdtd.py`` 1e-9, FP32 engine 1e-3; central
  differences (δε = 1e-3, FP64 engine) on 4 cells 1e-5 of the largest
  gradient;
* checkpoint length 1, 13 and N give bit-identical FP32 gradients;
* the torch.autograd.Function gives the same gradient as ``value_and_grad``.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import adjoint_oracle as ao
from oracle import fdtd_oracle as fo

pytestmark = pytest.mark.gpu

DL = 1e-3


def adj():
    import hipfdtd.adjoint

    return hipfdtd.adjoint


def pne(a, b):
    return fo.peak_normalised_error(a, b)


# ── small grid against the oracle's adjoint ──

SHAPE = (14, 15, 16)
NPML = 3
F0 = 3.0e10
BOX = (4, 10, 5, 11, 6, 12)


def wf(t):
    return fo.gaussian_sine(t, F0, F0)


SOURCES = [("x", (6, 7, 7, 8, 8, 9), wf), ("z", (3, 4, 4, 5, 3, 4), lambda t: 0.5 * wf(t))]
PROBES = [(0, 8, 8, 13), (0, 9, 8, 13), (2, 8, 9, 13)]
FREQS = [F0, 0.8 * F0]
STEPS = 70


def small_eps(seed=2):
    rng = np.random.default_rng(seed)
    eps = np.ones(SHAPE)
    eps[2:12, 3:13, 4:14] = 1.5
    i0, i1, j0, j1, k0, k1 = BOX
    design = rng.uniform(1.0, 4.0, (i1 - i0, j1 - j0, k1 - k0))
    return eps, design


def full(eps_bg, design, box):
    e = eps_bg.copy()
    i0, i1, j0, j1, k0, k1 = box
    e[i0:i1, j0:j1, k0:k1] = design
    return e


def oracle_small(eps_full):
    om = [2 * np.pi * f for f in FREQS]
    return ao.probe_dft_gradient(SHAPE, DL, NPML, eps_full, SOURCES, PROBES, om, STEPS,
                                 return_state=True)  # fmt: skip


@pytest.fixture(scope="module")
def small_oracle():
    eps_bg, design = small_eps()
    return eps_bg, design, oracle_small(full(eps_bg, design, BOX))


def box_of(vol, box):
    i0, i1, j0, j1, k0, k1 = box
    return vol[i0:i1, j0:j1, k0:k1]


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-9), (np.float32, 1e-4)])
def test_gradient_matches_oracle_adjoint(small_oracle, dtype, tol):
    eps_bg, design, (F_o, g_o, lam_o) = small_oracle
    print(f"[adjoint-oracle] dtype={np.dtype(dtype).name}", flush=True)
    A = adj()
    pb = A.ProbeDFTAdjoint(SHAPE, DL, NPML, BOX, SOURCES, PROBES, FREQS, STEPS,
                           eps_background=eps_bg, checkpoint=16, dtype=dtype)  # fmt: skip
    F = pb.forward(design)
    g = pb.backward(2.0 * F)
    ef = pne(F, F_o)
    eg = pne(g, box_of(g_o, BOX))
    print(f"[adjoint-oracle] F err {ef:.3e}  grad err {eg:.3e}", flush=True)
    assert ef < (1e-12 if dtype == np.float64 else 1e-5)
    assert eg < tol
    if dtype == np.float64:
        for c, name in enumerate(("Ex", "Ey", "Ez")):
            assert pne(pb.read_adjoint_volume(name), lam_o.E[c]) < 1e-9, name
        for c, name in enumerate(("Hx", "Hy", "Hz")):
            assert pne(pb.read_adjoint_volume(name), lam_o.H[c]) < 1e-9, name
    pb.close()


def test_checkpoint_length_does_not_change_the_gradient():
    eps_bg, design = small_eps(5)
    A = adj()
    grads = []
    for K in (1, 13, STEPS):
        print(f"[adjoint-ckpt] K={K}", flush=True)
        pb = A.ProbeDFTAdjoint(SHAPE, DL, NPML, BOX, SOURCES, PROBES, FREQS, STEPS,
                               eps_background=eps_bg, checkpoint=K)  # fmt: skip
        F = pb.forward(design)
        g = pb.backward(2.0 * F)
        # a second backward of the same forward recomputes the last segment too
        g2 = pb.backward(2.0 * F)
        assert np.array_equal(g, g2)
        grads.append(g)
        pb.close()
    assert np.array_equal(grads[0], grads[1])
    assert np.array_equal(grads[0], grads[2])


def test_forward_path_does_not_change_the_gradient():
    """The forward and recompute steps on the split, fused and march paths
    give bit-identical fields, so the gradient is the same on each."""
    eps_bg, design = small_eps(6)
    A = adj()
    grads = {}
    for path in ("split", "fused", "march"):
        print(f"[adjoint-path] {path}", flush=True)
        pb = A.ProbeDFTAdjoint(SHAPE, DL, NPML, BOX, SOURCES, PROBES, FREQS, STEPS,
                               eps_background=eps_bg, checkpoint=7)  # fmt: skip
        pb.fdtd.set_path(path)
        F = pb.forward(design)
        grads[path] = (F, pb.backward(2.0 * F))
        pb.close()
    for path in ("fused", "march"):
        np.testing.assert_array_equal(grads[path][0], grads["split"][0], err_msg=path)
        np.testing.assert_array_equal(grads[path][1], grads["split"][1], err_msg=path)


def test_autograd_function_matches_value_and_grad():
    import torch

    eps_bg, design = small_eps(7)
    A = adj()
    pb = A.ProbeDFTAdjoint(SHAPE, DL, NPML, BOX, SOURCES, PROBES, FREQS, STEPS,
                           eps_background=eps_bg, dtype=np.float64)  # fmt: skip
    J_ref, g_ref = pb.value_and_grad(design)
    t = torch.tensor(design, dtype=torch.float64, requires_grad=True)
    F = A.probe_dft(pb, t)
    J = (F.abs() ** 2).sum()
    J.backward()
    assert abs(float(J) - J_ref) <= 1e-12 * J_ref
    assert pne(t.grad.numpy(), g_ref) < 1e-12
    # a weighted objective: L = |F_0|² − 0.5 Re F_1
    t.grad = None
    F = A.probe_dft(pb, t)
    L = (F[0].abs() ** 2).sum() - 0.5 * F[1].real.sum()
    L.backward()
    gF = np.zeros(F.shape, dtype=np.complex128)
    Fn = F.detach().numpy()
    gF[0] = 2 * Fn[0]
    gF[1] = -0.5
    pb.forward(design)
    assert pne(t.grad.numpy(), pb.backward(gF)) < 1e-12
    pb.close()


def test_adjoint_argument_errors():
    A = adj()
    with pytest.raises(ValueError):
        A.ProbeDFTAdjoint(SHAPE, DL, NPML, (4, 20, 5, 11, 6, 12), SOURCES, PROBES, FREQS, 10)
    with pytest.raises(ValueError):
        A.ProbeDFTAdjoint(SHAPE, DL, NPML, BOX, SOURCES, PROBES + [PROBES[0]], FREQS, 10)
    pb = A.ProbeDFTAdjoint(SHAPE, DL, NPML, BOX, SOURCES, PROBES, FREQS, 10)
    with pytest.raises(RuntimeError):
        pb.backward(np.zeros((3, 2)))
    with pytest.raises(ValueError):
        pb.forward(np.ones((2, 2, 2)))
    pb.close()


# ── 40³ against torch FP64 autograd and central differences ──

S40 = (40, 40, 40)
NPML40 = 6
F40 = 2.5e10
BOX40 = (14, 26, 14, 26, 16, 24)
STEPS40 = 200


def wf40(t):
    return fo.gaussian_sine(t, F40, F40)


SRC40 = [("x", (20, 21, 20, 21, 10, 11), wf40)]
PROBES40 = [(0, i, j, 30) for i in (19, 20) for j in (19, 20)]


def design40():
    rng = np.random.default_rng(5)
    i0, i1, j0, j1, k0, k1 = BOX40
    return rng.uniform(1.0, 12.0, (i1 - i0, j1 - j0, k1 - k0))


@pytest.fixture(scope="module")
def torch40():
    import torch
    from torch_fdtd import probe_dft

    dev = "cuda" if torch.cuda.is_available() else "cpu"
    eps = torch.tensor(full(np.ones(S40), design40(), BOX40), dtype=torch.float64, device=dev,
                       requires_grad=True)  # fmt: skip
    F = probe_dft(eps, S40, DL, NPML40, SRC40, PROBES40, [2 * np.pi * F40], STEPS40)
    J = (F.abs() ** 2).sum()
    J.backward()
    return float(J), box_of(eps.grad.cpu().numpy(), BOX40)


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-9), (np.float32, 1e-3)])
def test_40cube_gradient_matches_torch_autograd(torch40, dtype, tol):
    J_t, g_t = torch40
    print(f"[adjoint-40] dtype={np.dtype(dtype).name}", flush=True)
    pb = adj().ProbeDFTAdjoint(S40, DL, NPML40, BOX40, SRC40, PROBES40, [F40], STEPS40,
                               checkpoint=50, dtype=dtype)  # fmt: skip
    J, g = pb.value_and_grad(design40())
    err = pne(g, g_t)
    print(f"[adjoint-40] J rel {abs(J - J_t) / J_t:.3e}  grad err {err:.3e}", flush=True)
    assert abs(J - J_t) <= (1e-12 if dtype == np.float64 else 1e-4) * J_t
    assert err < tol
    pb.close()


def test_40cube_gradient_matches_central_differences(torch40):
    _, g_t = torch40
    pb = adj().ProbeDFTAdjoint(S40, DL, NPML40, BOX40, SRC40, PROBES40, [F40], STEPS40,
                               checkpoint=50, dtype=np.float64)  # fmt: skip
    d0 = design40()
    J, g = pb.value_and_grad(d0)
    flat = np.argsort(np.abs(g).ravel())
    cells = [np.unravel_index(q, g.shape) for q in (flat[-1], flat[-7], flat[len(flat) // 2], flat[3])]
    d = 1e-3
    for c in cells:
        js = []
        for s in (+1, -1):
            d2 = d0.copy()
            d2[c] += s * d
            js.append(float(np.sum(np.abs(pb.forward(d2)) ** 2)))
        fd = (js[0] - js[1]) / (2 * d)
        print(f"[adjoint-fd] cell {c}: fd {fd:.6e} adjoint {g[c]:.6e}", flush=True)
        assert abs(fd - g[c]) <= 1e-5 * np.max(np.abs(g))
    pb.close()
