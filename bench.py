"""Benchmark of the HIP FDTD hot path on one MI355X.

Workload (the 256^3 FP32 configuration): 256^3 cells, dl = 1 mm, npml = 12;
εr from 40 random axis-aligned dielectric boxes (sides 8-48 cells, εr in
[1, 4], numpy seed 2) on vacuum; a 32x32 Jx sheet at z = 128 over
i, j in [112, 144) with rim taper 0.8 (renormalised), Gaussian-modulated sine
at f0 = c0/(20 dl), 50 % bandwidth; a Huygens box [32, 223]^3 recording the
DFT at 8 frequencies; after the run, the near-to-far transform at
181 θ x 2 φ angles, R = 1 m.  Inputs are synthetic (generated here).

A "step" is one full Yee time step on the FP32 step path the engine's
autotuner picks for this grid before anything is timed (``--path`` forces
one: ``split``, the H and E sweeps with the sheet source between them;
``fused`` or ``march``, one H + E sweep and the source fix-up; ``pair``,
one sweep for every two steps, the face DFT of the step between them read
from the cells the sweep records), and the
8-frequency face-DFT accumulation.  ``value`` is Mcell-updates/s over the K timed steps.
``roofline`` is the step's algorithmic HBM traffic (72 B/cell + 8 B per psi
entry) divided by the time of the step's sweep kernels, taken with HIP
events on the engine's stream around every sweep launch of the timed steps.  ``cpu_baseline``
times this repository's NumPy restatement of the reference's face-CPML engine
(one core) on a bounded sample of steps of the same grid.

The timed window ends with the face-DFT buffer folded into the sums (the
K = 32 matrix-core contractions of the window's samples all run inside it),
and after it the near-to-far transform runs at all 8 frequencies
(``n2f_ms``).  ``fp64_check`` runs the whole workload (``--accuracy-steps``,
default the configuration's 2000 steps) once more on the same FP32 path and
on the FP64 engine from the same inputs, and reports the peak-normalised
error of every field, every box-DFT component and the far E and H against
the FP64 run.

``roofline.traffic`` is the measured HBM traffic of the step's sweep
kernels, taken from ``profiles/split_sweeps_256cube_hbm_traffic.txt``
(``k_fused_256cube_hbm_traffic.txt``, ``k_march_256cube_hbm_traffic.txt``,
``k_pair_256cube_hbm_traffic.txt`` for the one-launch paths): two
rocprofv3 PMC passes, FETCH_SIZE and WRITE_SIZE, each scaled by what the same
counter reads for a plain device copy of known size (``--copy-calibration``
runs that copy).

``--workload adjoint`` measures the 160^3 inverse-design gradient: a
64x64x32 design box (εr ~ U[1, 12], seed 5), a Jx point source and a 4x4 Ex
probe patch 40 cells away along +z, J = Σ|DFT(Ex)|² at one frequency, 3000
steps with a checkpoint every 100.  It reports the adjoint sweep's wall time
over the forward run's (recomputation of the checkpointed segments plus the
transposed steps) and checks the FP32 gradient on 16 random design cells
against central differences (δε = 1e-3) of FP64 runs.

``--workload batch`` measures 64 simulations of uneven sizes (sides
U{48..160}, seed 4, npml 8, a centred dielectric sphere, a centred Jz dipole
and a 4-frequency Ez probe each, 1200 steps) advanced together by the
batched sweep (path and schedule picked by the autotuner among the split,
fused and march batch paths; HIP-graph replays), against the same 64 runs
one after another on the plain engine: with its device point probe, on the
fused sweep, and through the reference's plugin API alone (source callback,
host DFT monitor reading the probe cell each step).

``--workload plane-dft`` (not the default line) measures the 512^3 plane-DFT
configuration: smooth Gaussian random εr (seed 3, σ = 8 cells, clipped to
[1, 6]), three orthogonal point dipoles with staggered pulses, npml 16, and a
16-frequency DFT of Ex, Ey, Ez on the three mid-planes.  It reports the
per-step cost of each DFT mode (K = 32 buffered matrix-core contraction,
direct FP32, direct Kahan, direct FP64) and, after the full run, each mode's
error against the FP64 sums.

Usage: python bench.py [--gpus 1] [--steps 2000] [--warmup 20] [--no-cpu]
       python bench.py --copy-calibration
       python bench.py --workload stream-ceiling
       python bench.py --workload plane-dft [--steps 1500]
       python bench.py --workload adjoint [--steps 3000] [--repeat 2]
       python bench.py --workload batch [--steps 1200] [--graph-steps 100]
                       [--batch-path auto|split|fused|march] [--no-serial]
       python bench.py --workload reference-sizes
"""

from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "opencl-fdtd-solver_amd"))

C0 = 299_792_458.0
HBM_PEAK_GBS = 8000.0
TIMING_EVERY = 15  # one sweep launch in 15 is timed with HIP events
FIELDS = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")


def c2_epsilon(shape, npml, seed=2, n_boxes=40):
    """40 random dielectric boxes inside [npml, N - npml)."""
    rng = np.random.default_rng(seed)
    eps = np.ones(shape, dtype=np.float32)
    for _ in range(n_boxes):
        size = rng.integers(8, 49, size=3)
        lo = [int(rng.integers(npml, shape[a] - npml - size[a] + 1)) for a in range(3)]
        er = float(rng.uniform(1.0, 4.0))
        eps[lo[0] : lo[0] + size[0], lo[1] : lo[1] + size[1], lo[2] : lo[2] + size[2]] = er
    return eps


def gaussian_sine(t, f0, fw):
    t0 = 5.0 / (math.pi * fw)
    sig = 1.0 / (math.pi * fw)
    return math.exp(-0.5 * ((t - t0) / sig) ** 2) * math.sin(2 * math.pi * f0 * t)


class Workload:
    shape = (256, 256, 256)
    npml = 12
    dl = 1e-3
    box = (32, 223)
    n_freq = 8

    def __init__(self):
        self.f0 = C0 / (20 * self.dl)
        self.fw = 0.5 * self.f0
        self.freqs = np.linspace(self.f0 - self.fw / 2, self.f0 + self.fw / 2, self.n_freq)
        self.eps = c2_epsilon(self.shape, self.npml)
        lo, hi = self.box
        self.center = ((lo + hi) / 2 * self.dl,) * 3
        self.size = ((hi - lo) * self.dl,) * 3

    def source(self, f):
        f.add_source_Jx(128, gaussian_sine(f.t, self.f0, self.fw), 112, 144, 112, 144,
                        rim_taper=True, rim_edge=0.8)  # fmt: skip

    def psi_entries(self):
        nx, ny, nz = self.shape
        n = self.npml
        return 4 * (2 * n * ny * nz + nx * 2 * n * nz + nx * ny * 2 * n)

    def obs_points(self, R=1.0):
        th = np.deg2rad(np.linspace(0.0, 180.0, 181))
        pts = []
        for ph in (0.0, math.pi / 2):
            pts.append(np.column_stack([R * np.sin(th) * math.cos(ph),
                                        R * np.sin(th) * math.sin(ph), R * np.cos(th)]))  # fmt: skip
        return np.concatenate(pts)


def run_gpu(args, wl: Workload):
    import hipfdtd

    path = None if args.path == "auto" else args.path
    sim = hipfdtd.HIPFDTD(wl.shape, wl.dl, npml=wl.npml, dtype=np.float32, path=path)
    tuned = None
    if args.path == "march" and args.march_tiling:
        v = [int(x) for x in args.march_tiling.split(",")]
        sim.set_march_tiling(v[0], v[1], v[2] if len(v) > 2 else 0, v[3] if len(v) > 3 else -2)
    sim.set_epsilon(wl.eps)
    if args.path == "auto":
        # the engine's autotuner picks the step path and launch shape (on a
        # scratch engine with the same kind of coefficients, before anything
        # is timed)
        from hipfdtd.autotune import tune_engine

        tuned = tune_engine(sim, rows=(8,), planes=(-1,), march_rows=(1, 2),
                            march_planes=(-1, 24, 40), steps=10, warmup=2, use_cache=False,
                            write_cache=False)  # fmt: skip
        print(f"[bench] autotuned path {tuned['path']} rows={tuned['rows']} "
              f"planes={tuned['planes']}: {tuned['table']}", file=sys.stderr, flush=True)
    sim.add_source(wl.source)
    mon = hipfdtd.HIPNear2FarMonitor(sim, wl.center, wl.size, freqs=wl.freqs)
    assert (mon.ix0, mon.ix1) == wl.box, (mon.ix0, mon.ix1)
    sim.run(args.warmup)  # the pair path takes two steps per launch in run()
    mon.flush()  # the window starts with an empty K-step buffer
    sim.synchronize()
    # HIP events on a sample of the window's sweep launches (every 15th: an
    # odd stride times the split path's H and E sweeps alike), recorded by
    # each sampled launch around its kernel; a timed launch costs the stream
    # a few microseconds, so timing every launch would slow the window it
    # measures
    sim.kernel_timing(True, every=TIMING_EVERY)
    t0 = time.perf_counter()
    sim.run(args.steps)
    # every face-DFT sample of the window is folded into the sums inside it:
    # the K = 32 contractions of full blocks and one for the partial block
    mon.flush()
    sim.synchronize()
    t1 = time.perf_counter()
    kern_ms, timed = sim.kernel_timing_read()
    launches = sim.kernel_timing_launches()
    sim.kernel_timing(False)
    print(f"[bench] timed {args.steps} steps in {t1 - t0:.3f} s", file=sys.stderr, flush=True)

    obs = wl.obs_points()
    t2 = time.perf_counter()
    eh = [mon.get_farfields(obs, freq_index=f) for f in range(wl.n_freq)]
    t3 = time.perf_counter()
    finite = bool(np.all(np.isfinite(eh)))
    return dict(wall_s=t1 - t0, kern_ms=kern_ms, timed=timed, launches=launches,
                n2f_ms=(t3 - t2) * 1e3,
                finite=finite, device=sim.device_name, path=sim.path,
                tuned=tuned)  # fmt: skip


def run_accuracy(wl: Workload, tuned, steps: int, path=None):
    """The whole workload (``steps`` steps, box DFT, far field) on the FP32
    engine the benchmark times (same path and launch shape) and on the FP64
    engine, from the same inputs.  Peak-normalised errors of the FP32 run
    against the FP64 one: the E and H fields after the last step (each group
    normalised by its peak), the E and H components of the box DFT (worst of
    the 8 frequencies), and the far E and H at 8 frequencies x 362 angles
    (worst frequency); ``per_component`` normalises every field and DFT
    component by its own peak instead."""
    import hipfdtd
    from hipfdtd.autotune import apply

    obs = wl.obs_points()
    out = {}
    for dtype in (np.float32, np.float64):
        sim = hipfdtd.HIPFDTD(wl.shape, wl.dl, npml=wl.npml, dtype=dtype,
                              path=path if dtype == np.float32 else None)
        sim.set_epsilon(wl.eps)
        if dtype == np.float32 and tuned is not None:
            apply(sim, tuned)
        sim.add_source(wl.source)
        mon = hipfdtd.HIPNear2FarMonitor(sim, wl.center, wl.size, freqs=wl.freqs)
        sim.run(steps)
        fields = sim.fields()
        dft = {c: np.stack([mon.read_dft(c, f) for f in range(wl.n_freq)]) for c in FIELDS}
        ff = np.stack([mon.get_farfields(obs, freq_index=f) for f in range(wl.n_freq)])
        ffp = np.stack([mon.get_farfields(obs, freq_index=f, precise=True)
                        for f in range(wl.n_freq)])  # fmt: skip
        out[np.dtype(dtype).name] = (fields, dft, (ff, ffp), sim.path)
        sim.close()
    (f32, d32, (ff32, ffp32), path32) = out["float32"]
    (f64, d64, (ff64, _), path64) = out["float64"]

    def pn(a, b):
        den = float(np.max(np.abs(b)))
        return float(np.max(np.abs(a.astype(b.dtype) - b))) / den if den > 0 else 0.0

    def pn_group(a, b, comps):
        den = max(float(np.max(np.abs(b[c]))) for c in comps)
        num = max(float(np.max(np.abs(a[c].astype(b[c].dtype) - b[c]))) for c in comps)
        return num / den if den > 0 else 0.0

    groups = {"E": FIELDS[:3], "H": FIELDS[3:]}
    fields = {c: pn(f32[c], f64[c]) for c in FIELDS}
    box = {c: max(pn(d32[c][f], d64[c][f]) for f in range(wl.n_freq)) for c in FIELDS}
    fields_g = {g: pn_group(f32, f64, cs) for g, cs in groups.items()}
    box_g = {g: max(pn_group({c: d32[c][f] for c in cs}, {c: d64[c][f] for c in cs}, cs)
                    for f in range(wl.n_freq)) for g, cs in groups.items()}  # fmt: skip
    def far_err(a):
        return {"E": max(pn(a[f][:, :3], ff64[f][:, :3]) for f in range(wl.n_freq)),
                "H": max(pn(a[f][:, 3:], ff64[f][:, 3:]) for f in range(wl.n_freq))}

    # the reference rounds the observation points and k to the field dtype
    # and forms k r.r' in it (monitors.py:653, 669); "farfield_fp64_geometry"
    # keeps those in FP64 and leaves only the FP32 fields and sums
    far = far_err(ff32)
    return {"steps": steps, "reference": "FP64 HIP engine, same inputs (path %s)" % path64,
            "fp32_path": path32,
            "per_component": {"fields": fields, "box_dft": box},
            "fields": fields_g, "box_dft": box_g, "farfield": far,
            "farfield_fp64_geometry": far_err(ffp32),
            "max": max(list(fields_g.values()) + list(box_g.values()) + list(far.values()))}  # fmt: skip


def run_cpu_baseline(wl: Workload, steps: int, warmup: int = 2):
    """NumPy face-CPML restatement (one core) on `steps` steps of the same
    grid after `warmup` untimed ones."""
    from oracle import fdtd_oracle as fo

    sim = fo.OracleFDTD(wl.shape, wl.dl, npml=wl.npml)
    sim.set_epsilon(wl.eps)
    sim.add_source(wl.source)
    for _ in range(warmup):
        sim.step()
    t0 = time.perf_counter()
    for _ in range(steps):
        sim.step()
    dt = time.perf_counter() - t0
    cells = float(np.prod(wl.shape))
    return cells * steps / dt / 1e6, dt


def c3_epsilon(shape, seed=3, sigma=8.0):
    """Smooth Gaussian random field: white noise low-passed by a Gaussian of
    `sigma` cells (FFT), standardised; εr = clip(3.5 + 1.25 g, 1, 6)."""
    import scipy.fft as sf

    rng = np.random.default_rng(seed)
    g = rng.standard_normal(shape, dtype=np.float32)
    spec = sf.rfftn(g, workers=-1)
    kx = sf.fftfreq(shape[0]).astype(np.float32)[:, None, None]
    ky = sf.fftfreq(shape[1]).astype(np.float32)[None, :, None]
    kz = sf.rfftfreq(shape[2]).astype(np.float32)[None, None, :]
    c = np.float32(-2.0 * (math.pi * sigma) ** 2)
    spec *= np.exp(c * (kx * kx + ky * ky + kz * kz))
    g = sf.irfftn(spec, s=shape, workers=-1).astype(np.float32)
    g = (g - g.mean()) / g.std()
    return np.clip(3.5 + 1.25 * g, 1.0, 6.0).astype(np.float32)


def run_plane_dft(args):
    """512^3 16-frequency mid-plane DFT: cost per mode and error vs FP64."""
    import hipfdtd
    import hipfdtd.dft

    shape, npml, dl = (512, 512, 512), 16, 1e-3
    f0 = C0 / (20 * dl)
    fw = 0.5 * f0
    freqs = np.linspace(f0 - fw / 2, f0 + fw / 2, 16)
    eps = c3_epsilon(shape)
    rng = np.random.default_rng(3)
    pos = [tuple(int(rng.integers(npml + 16, s - npml - 16)) for s in shape) for _ in range(3)]

    def make():
        sim = hipfdtd.HIPFDTD(shape, dl, npml=npml)
        sim.set_epsilon(eps)
        for comp, p, lag in zip("xyz", pos, (0, 20, 40)):
            box = (p[0], p[0] + 1, p[1], p[1] + 1, p[2], p[2] + 1)
            sim.add_source(lambda f, comp=comp, box=box, lag=lag:
                           f.add_source_J(comp, gaussian_sine(f.t - lag * f.dt, f0, fw), box))  # fmt: skip
        return sim

    modes = {"buffered_k32": dict(mode="buffered", K=32), "direct_fp32": dict(mode="direct"),
             "direct_kahan": dict(mode="direct", accum="kahan"),
             "direct_fp64": dict(mode="direct", accum="fp64")}  # fmt: skip

    def attach(sim, kw):
        return [hipfdtd.dft.HIPPlaneDFTMonitor(sim, ax, 256, ("Ex", "Ey", "Ez"), freqs, **kw)
                for ax in "xyz"]  # fmt: skip

    def timed(sim, n):
        sim.synchronize()
        t0 = time.perf_counter()
        sim.run(n)
        sim.synchronize()
        return (time.perf_counter() - t0) / n * 1e6

    sim = make()
    sim.run(32)
    base = timed(sim, 96)
    cost = {}
    for name, kw in modes.items():
        mons = attach(sim, kw)
        sim.run(32)
        cost[name] = round(timed(sim, 96) - base, 1)
        for m in mons:
            sim._monitors.remove(m)
            m._release()
    sim.close()
    print(f"[plane-dft] per-step cost {cost}, sweep {base:.1f} us", file=sys.stderr, flush=True)

    sim = make()
    allm = {name: attach(sim, kw) for name, kw in modes.items()}
    sim.run(args.steps)
    ref = [m.dft() for m in allm["direct_fp64"]]
    err = {}
    for name in ("buffered_k32", "direct_fp32", "direct_kahan"):
        # planes the pulse has not reached yet hold exact zeros in every mode
        errs = [float(np.max(np.abs(m.dft() - r)) / np.max(np.abs(r)))
                for m, r in zip(allm[name], ref) if np.max(np.abs(r)) > 0]  # fmt: skip
        err[name] = max(errs) if errs else None
    print(json.dumps({
        "metric": "plane_dft_us_per_step", "unit": "us", "higher_is_better": False,
        "steps": args.steps, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "512^3 FP32 npml 16, smooth random eps in [1,6], 3 dipoles; "
                               "16-frequency DFT of Ex,Ey,Ez on the 3 mid-planes",
                   "samples_per_step": 3 * 3 * 512 * 512},
        "sweep_us_per_step": round(base, 1),
        "dft_us_per_step": cost,
        "max_rel_err_vs_fp64": err,
        "device": sim.device_name,
    }))  # fmt: skip


def run_adjoint(args):
    """160^3 inverse-design gradient: forward and checkpointed adjoint sweep
    timed on the FP32 engine, gradient checked against central differences of
    FP64 runs on 16 random design cells."""
    from hipfdtd.adjoint import ProbeDFTAdjoint

    shape, npml, dl = (160, 160, 160), 12, 1e-3
    f0 = C0 / (30 * dl)
    fw = 0.5 * f0
    rng = np.random.default_rng(5)
    box = (48, 112, 48, 112, 64, 96)
    design = rng.uniform(1.0, 12.0, (64, 64, 32))
    src = [("x", (80, 81, 80, 81, 58, 59), lambda t: gaussian_sine(t, f0, fw))]
    probes = [(0, i, j, 98) for i in range(78, 82) for j in range(78, 82)]
    steps, K = args.steps, 100
    pb = ProbeDFTAdjoint(shape, dl, npml, box, src, probes, [f0], steps, checkpoint=K)
    for _ in range(args.warmup):
        pb.value_and_grad(design)
    tf, tb, trc, trv = [], [], [], []
    for _ in range(args.repeat):
        F = pb.forward(design)
        g = pb.backward(2.0 * F)
        tf.append(pb.timing["forward_s"])
        tb.append(pb.timing["backward_s"])
        trc.append(pb.timing["recompute_s"])
        trv.append(pb.timing["reverse_s"])
    J = float(np.sum(np.abs(F) ** 2))
    mem32 = pb.memory_bytes()
    dev = pb.fdtd.device_name
    pb.close()
    t_f, t_b = float(np.median(tf)), float(np.median(tb))
    print(f"[adjoint] forward {t_f:.3f} s backward {t_b:.3f} s", file=sys.stderr, flush=True)

    p64 = ProbeDFTAdjoint(shape, dl, npml, box, src, probes, [f0], steps, checkpoint=K,
                          dtype=np.float64)  # fmt: skip
    J64, g64 = p64.value_and_grad(design)
    cells = [tuple(int(rng.integers(0, n)) for n in design.shape) for _ in range(16)]
    d = 1e-3
    fd = []
    for c in cells:
        js = []
        for sgn in (+1, -1):
            d2 = design.copy()
            d2[c] += sgn * d
            js.append(float(np.sum(np.abs(p64.forward(d2)) ** 2)))
        fd.append((js[0] - js[1]) / (2 * d))
    p64.close()
    fd = np.array(fd)
    a32 = np.array([g[c] for c in cells])
    a64 = np.array([g64[c] for c in cells])
    cells_n = float(np.prod(shape))
    print(json.dumps({
        "metric": "adjoint/forward time", "value": round(t_b / t_f, 3), "unit": "x",
        "higher_is_better": False, "steps": steps, "warmup": args.warmup,
        "repeat": args.repeat, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "160^3 FP32 npml 12, design box 64x64x32 centred with eps_r "
                               "~ U[1,12] (seed 5), Jx point source, 4x4 Ex probe patch 40 "
                               "cells away along +z, J = sum |DFT(Ex)|^2 at f0 = c0/(30 dl)",
                   "checkpoint_every": K, "design_cells": int(design.size)},
        "forward_s": round(t_f, 4), "backward_s": round(t_b, 4),
        "backward_recompute_s": round(float(np.median(trc)), 4),
        "backward_reverse_s": round(float(np.median(trv)), 4),
        "forward_mcell_per_s": round(cells_n * steps / t_f / 1e6, 1),
        "reverse_mcell_per_s": round(cells_n * steps / float(np.median(trv)) / 1e6, 1),
        "device_bytes_fp32": mem32,
        "objective_fp32_vs_fp64_rel": abs(J - J64) / J64,
        "fd_check": {"cells": 16, "delta_eps": d, "fd_engine": "f64",
                     "max_rel_err_f32_adjoint": float(np.max(np.abs(a32 - fd) / np.abs(fd))),
                     "max_rel_err_f64_adjoint": float(np.max(np.abs(a64 - fd) / np.abs(fd))),
                     "peak_norm_err_f32_adjoint": float(np.max(np.abs(a32 - fd)) / np.max(np.abs(fd)))},
        "grad_f32_vs_f64_peak_norm": float(np.max(np.abs(g - g64)) / np.max(np.abs(g64))),
        "device": dev,
    }))  # fmt: skip


def c4_sims(n=64, seed=4, dl=1e-3, npml=8):
    """64 uneven grids: sides U{48..160}, a centred sphere (radius U[5, 30]
    cells, clipped to fit; εr U[1.5, 8]), a centred Jz dipole at
    f ~ U(c0/(30 dl), c0/(12 dl)), an Ez probe beside the sphere, 4 DFT
    frequencies around f."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        shape = tuple(int(v) for v in rng.integers(48, 161, 3))
        r = float(rng.uniform(5.0, 30.0))
        er = float(rng.uniform(1.5, 8.0))
        f = float(rng.uniform(C0 / (30 * dl), C0 / (12 * dl)))
        c = [v // 2 for v in shape]
        r = min(r, min(shape) // 2 - npml - 2)
        pi = min(c[0] + int(r) + 3, shape[0] - npml - 2)
        out.append({"shape": shape, "r": r, "er": er, "f": f, "centre": c,
                    "probe": (pi, c[1], c[2]), "freqs": list(f * np.array([0.8, 14 / 15, 16 / 15, 1.2]))})  # fmt: skip
    return out, dl, npml


def c4_eps(sim):
    nx, ny, nz = sim["shape"]
    I, J, K = np.ogrid[:nx, :ny, :nz]
    c = sim["centre"]
    eps = np.ones(sim["shape"], dtype=np.float32)
    eps[(I - c[0]) ** 2 + (J - c[1]) ** 2 + (K - c[2]) ** 2 < sim["r"] ** 2] = sim["er"]
    return eps


def run_batch(args):
    """64 uneven grids: one persistent batched launch per step (graph-replayed)
    against the same 64 runs one after another on the plain engine."""
    import hipfdtd
    from hipfdtd.autotune import tune_batch
    from hipfdtd.batch import BatchedFDTD
    from hipfdtd.probes import HIPPointProbe

    cfg, dl, npml = c4_sims()
    steps = args.steps
    cells = float(sum(np.prod(c["shape"]) for c in cfg))
    psi = float(sum(4 * 2 * npml * (c["shape"][1] * c["shape"][2] + c["shape"][0] * c["shape"][2]
                                     + c["shape"][0] * c["shape"][1]) for c in cfg))  # fmt: skip
    epss = [c4_eps(c) for c in cfg]

    def wave(c):
        f = c["f"]
        return lambda t: gaussian_sine(t, f, 0.5 * f)

    def box(c):
        x, y, z = c["centre"]
        return (x, x + 1, y, y + 1, z, z + 1)

    def serial_run(fused):
        """The plain engine, one grid after another, source and probe as callbacks."""
        total, dfts, dev = 0.0, [], ""
        for c, eps in zip(cfg, epss):
            sim = hipfdtd.HIPFDTD(c["shape"], dl, npml=npml, fused=fused)
            dev = sim.device_name
            sim.set_epsilon(eps)
            w, bx = wave(c), box(c)
            sim.add_source(lambda f, w=w, bx=bx: f.add_source_J("z", w(f.t), bx))
            pr = HIPPointProbe(sim, c["probe"], "Ez", c["freqs"])
            sim.synchronize()
            t0 = time.perf_counter()
            sim.run(steps)
            sim.synchronize()
            total += time.perf_counter() - t0
            dfts.append(pr.dft())
            sim.close()
        return total, dfts, dev

    def serial_plain_api():
        """The plain engine driven only through the reference's plugin API:
        the source as a callback, the probe DFT as a host monitor reading
        the cell with read_point after every step (the reference has no
        device-side probe)."""
        total, dfts = 0.0, []
        for c, eps in zip(cfg, epss):
            sim = hipfdtd.HIPFDTD(c["shape"], dl, npml=npml)
            sim.set_epsilon(eps)
            w, bx = wave(c), box(c)
            sim.add_source(lambda f, w=w, bx=bx: f.add_source_J("z", w(f.t), bx))
            om = 2 * np.pi * np.asarray(c["freqs"])
            acc = np.zeros(len(om), complex)
            cell = tuple(c["probe"])

            def probe(f, cell=cell, om=om, acc=acc):
                acc += float(f.read_point("Ez", *cell)) * np.exp(1j * om * f.t) * f.dt

            sim.add_monitor(probe)
            sim.synchronize()
            t0 = time.perf_counter()
            sim.run(steps)
            sim.synchronize()
            total += time.perf_counter() - t0
            dfts.append(acc)
            sim.close()
        return total, dfts

    # the plain engine's default path (split sweeps for FP32), and the fused sweep
    if args.no_serial:  # profiling runs: the batch alone
        t_serial = t_fused = t_plain = plain_err = None
        serial, dev = None, ""
    else:
        t_serial, serial, dev = serial_run(None)
        t_fused, _, _ = serial_run(True)
        t_plain, plain = serial_plain_api()
        plain_err = float(np.max(np.abs(np.array(plain) - np.array(serial))
                                 / np.max(np.abs(np.array(serial)), axis=1, keepdims=True)))  # fmt: skip
        print(f"[batch] serial {t_serial:.2f} s (fused sweep {t_fused:.2f} s, plugin API with a "
              f"host probe {t_plain:.2f} s)", file=sys.stderr, flush=True)  # fmt: skip

    sims = []
    for c, eps in zip(cfg, epss):
        sim = hipfdtd.HIPFDTD(c["shape"], dl, npml=npml)
        sim.set_epsilon(eps)
        sims.append(sim)
    batch = BatchedFDTD(sims)
    paths = ("fused", "split", "march") if args.batch_path == "auto" else (args.batch_path,)
    tuned = tune_batch(batch, paths=paths, use_cache=False, write_cache=False)
    dev = dev or sims[0].device_name
    batch.set_sources([("z", box(c)) for c in cfg], [wave(c) for c in cfg], steps)
    batch.set_probes([c["probe"] for c in cfg], "Ez", [c["freqs"] for c in cfg], steps)
    sims[0].synchronize()
    t0 = time.perf_counter()
    batch.run(steps, graph_steps=args.graph_steps)
    t_batch = time.perf_counter() - t0
    got = batch.probe_dfts()
    err = None
    if serial is not None:
        ref = np.array(serial)
        err = float(np.max(np.abs(got - ref) / np.max(np.abs(ref), axis=1, keepdims=True)))
    batch.close()
    for s in sims:
        s.close()
    b_alg = 72.0 * cells + 8.0 * psi
    step_s = t_batch / steps
    print(json.dumps({
        "metric": "batched/serial throughput",
        "value": round(t_serial / t_batch, 2) if t_serial else None, "unit": "x",
        "higher_is_better": True, "steps": steps, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "64 grids, sides U{48..160} (seed 4), npml 8, centred sphere "
                               "r U[5,30] eps_r U[1.5,8], centred Jz dipole f ~ U(c0/30dl, "
                               "c0/12dl), Ez probe, 4-frequency DFT",
                   "total_cells": int(cells), "graph_steps": args.graph_steps},
        "batch_mcell_per_s": round(cells * steps / t_batch / 1e6, 1),
        "serial_mcell_per_s": round(cells * steps / t_serial / 1e6, 1) if t_serial else None,
        "batch_s": round(t_batch, 3), "serial_s": round(t_serial, 3) if t_serial else None,
        "serial_path": "default (split sweeps for FP32)",
        "serial_fused_sweep_s": round(t_fused, 3) if t_fused else None,
        "batched_vs_serial_fused_sweep": round(t_fused / t_batch, 2) if t_fused else None,
        "serial_plain_api_s": round(t_plain, 3) if t_plain else None,
        "serial_plain_api": "source callback + host DFT monitor reading the probe cell with "
                            "read_point each step (the reference's plugin API)",
        "batched_vs_serial_plain_api": round(t_plain / t_batch, 2) if t_plain else None,
        "plain_api_probe_max_rel_err_vs_device_probe": plain_err,
        "batch_ms_per_step": round(step_s * 1e3, 3),
        "batch_hbm_gbs_alg": round(b_alg / step_s / 1e9, 1),
        "batch_hbm_frac": round(b_alg / step_s / 1e9 / HBM_PEAK_GBS, 4),
        "probe_dft_max_rel_err_vs_serial": err,
        "autotune": {k: tuned[k] for k in ("path", "blocks", "rows", "planes", "width", "ms")},
        "autotune_table": tuned["table"],
        "device": dev,
    }))  # fmt: skip


def run_reference_sizes(args):
    """The reference's own throughput benchmark (benchmarks/benchmark_vs_meep.py:62-135,
    README.md:184-190) at its published sizes: vacuum, dl = 1 mm, npml 25, a
    soft Ex sheet over the whole xy plane at z = nz - npml - 2 driven by a
    6 GHz Gaussian-modulated sine (20 % bandwidth), no monitors; 10 warm-up
    steps, then the median of 3 timed windows."""
    import warnings

    import hipfdtd

    warnings.simplefilter("ignore", DeprecationWarning)
    cases = [((400, 400, 400), 100), ((600, 600, 600), 100), ((750, 750, 750), 80)]
    freq, npml, dl = 6e9, 25, 1e-3
    fw = 0.2 * freq
    t0, sig = 5.0 / (math.pi * fw), 1.0 / (math.pi * fw)
    out = []
    dev = ""
    for shape, steps in cases:
        sim = hipfdtd.HIPFDTD(shape, dl, npml=npml)
        dev = sim.device_name
        z_src = shape[2] - npml - 2

        def src(f, z=z_src):
            amp = math.exp(-0.5 * ((f.t - t0) / sig) ** 2) * math.sin(2 * math.pi * freq * f.t)
            f.add_source_Ex(z, amp)  # the reference's legacy soft sheet

        sim.add_source(src)
        sim.run(10)
        sim.synchronize()
        rates = []
        for _ in range(3):
            t_a = time.perf_counter()
            sim.run(steps)
            sim.synchronize()
            rates.append(float(np.prod(shape)) * steps / (time.perf_counter() - t_a) / 1e6)
        out.append({"grid": list(shape), "steps": steps, "path": sim.path,
                    "mcups_median": round(float(np.median(rates)), 1),
                    "mcups": [round(r, 1) for r in rates]})  # fmt: skip
        print(f"[reference-sizes] {shape} {sim.path}: {out[-1]['mcups_median']} Mcell/s",
              file=sys.stderr, flush=True)  # fmt: skip
        sim.close()
    print(json.dumps({
        "metric": "Mcell-updates/s", "unit": "Mcell/s", "higher_is_better": True,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": "reference throughput benchmark: vacuum, npml 25, soft Ex sheet "
                               "over the whole xy plane, no monitors; median of 3 windows"},
        "results": out,
        "reference_published_rtx5080": {"400^3": 7970, "600^3": 8630, "750^3": 9430},
        "device": dev,
    }))  # fmt: skip


TRAFFIC_FILES = {"fused": ROOT / "profiles" / "k_fused_256cube_hbm_traffic.txt",
                 "split": ROOT / "profiles" / "split_sweeps_256cube_hbm_traffic.txt",
                 "march": ROOT / "profiles" / "k_march_256cube_hbm_traffic.txt",
                 "pair": ROOT / "profiles" / "k_pair_256cube_hbm_traffic.txt"}  # fmt: skip

KERNELS = {"fused": "k_fused (fused H + E sweep; sources by k_source_fixup)",
           "split": "k_update_H + k_update_E (split sweeps, one step; sources by k_add_source)",
           "march": "k_march (register-pipelined H + E sweep, sources applied in the sweep)",
           "pair": "k_pair (two H + E steps per launch, sources applied in the sweep)"}


def measured_traffic(path):
    """HBM bytes per step of the sweep kernels recorded in the profile
    summary for this path, or None."""
    try:
        for line in TRAFFIC_FILES[path].read_text().splitlines():
            if line.startswith("traffic_bytes_per_step:"):
                return float(line.split(":", 1)[1])
    except (OSError, KeyError):
        pass
    return None


def copy_calibration(n_bytes=1 << 30, reps=5):
    """A plain device-to-device copy of n_bytes, reps times (the reference
    kernel for scaling the FETCH_SIZE / WRITE_SIZE counters)."""
    import torch

    src = torch.ones(n_bytes // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": n_bytes, "copy_bytes_written": n_bytes, "reps": reps}))


def stream_ceiling(n_bytes=1 << 30, reps=10):
    """What plain streaming kernels reach on this device: the library's
    stream probe (16-byte loads and stores, grid-stride loops) over 1 GiB
    arrays (256 MiB for the six-array kernel), at several grid sizes, the best
    of each kind.  copy = 1 read + 1 write stream, triad = 2 + 1, six = six
    arrays read and six written (the balance of a Yee step's field traffic),
    read = a sum.  It is the practical ceiling for a sweep's mix of reads
    and writes, below the 8 TB/s peak."""
    import ctypes as C

    from hipfdtd import _lib

    lib = _lib.load()
    kinds = {"copy": (0, n_bytes, 2.0), "triad": (1, n_bytes, 3.0),
             "six_by_six": (2, n_bytes // 4, 12.0), "read": (3, n_bytes, 1.0)}  # fmt: skip
    out = {}
    for name, (kind, nb, streams) in kinds.items():
        best = None
        for blocks in (256, 512, 1024, 2048, 4096):
            sec = C.c_double(0.0)
            _lib.check(lib.hfdtd_stream_probe(0, kind, nb, blocks, reps, C.byref(sec)),
                       "hfdtd_stream_probe")  # fmt: skip
            tbs = streams * nb / sec.value / 1e12
            if best is None or tbs > best["tb_per_s"]:
                best = {"tb_per_s": round(tbs, 3), "blocks": blocks,
                        "us": round(sec.value * 1e6, 1)}  # fmt: skip
        out[name] = best
    print(json.dumps({"stream_ceiling": out, "bytes_per_array": {k: v[1] for k, v in kinds.items()},
                      "reps": reps, "peak_tb_per_s": HBM_PEAK_GBS / 1e3}))  # fmt: skip


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--cpu-steps", type=int, default=20)
    ap.add_argument("--cpu-warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--accuracy-steps", type=int, default=2000,
                    help="256^3 workload: steps of the FP32-vs-FP64 accuracy run (0 = skip)")
    ap.add_argument("--copy-calibration", action="store_true")
    ap.add_argument("--workload", default="yee-256",
                    choices=["yee-256", "plane-dft", "adjoint", "batch", "reference-sizes",
                             "stream-ceiling"])
    ap.add_argument("--graph-steps", type=int, default=100, help="batch: steps per HIP graph")
    ap.add_argument("--batch-path", default="auto", choices=["auto", "split", "fused", "march"],
                    help="batch: tune only this path (auto: all three)")
    ap.add_argument("--no-serial", action="store_true",
                    help="batch: skip the serial runs (profiling the batch alone)")
    ap.add_argument("--repeat", type=int, default=2, help="adjoint: timed forward+backward runs")
    ap.add_argument("--march-tiling", default=None, metavar="ROWS,PLANES[,RUN[,NT]]",
                    help="with --path march: rows per wavefront, planes per block, blocks per XCD "
                         "run, non-temporal streams (1 / 0)")
    ap.add_argument("--path", default="auto", choices=["auto", "fused", "split", "march", "pair"],
                    help="256^3 workload: engine step path (auto = picked by the autotuner)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.workload == "stream-ceiling":
        stream_ceiling()
        return
    if args.workload == "plane-dft":
        if args.steps == 2000:
            args.steps = 1500
        run_plane_dft(args)
        return
    if args.workload == "batch":
        if args.steps == 2000:
            args.steps = 1200
        run_batch(args)
        return
    if args.workload == "reference-sizes":
        run_reference_sizes(args)
        return
    if args.workload == "adjoint":
        if args.steps == 2000:
            args.steps = 3000
        if args.warmup == 20:
            args.warmup = 1
        run_adjoint(args)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); the multi-GPU path "
              "(x-slab halo exchange) is design-only in this build.", file=sys.stderr)  # fmt: skip
        sys.exit(2)
    wl = Workload()
    res = run_gpu(args, wl)
    cells = float(np.prod(wl.shape))
    mcells = cells * args.steps / res["wall_s"] / 1e6
    ms_step = res["wall_s"] / args.steps * 1e3
    b_alg = 72.0 * cells + 8.0 * wl.psi_entries()
    # sweep kernels per step: k_fused, k_march, k_update_H + k_update_E, or
    # half a k_pair launch (mean of the timed launches x launches per step)
    step_kern_s = res["kern_ms"] / res["timed"] * res["launches"] / args.steps / 1e3
    achieved = b_alg / step_kern_s / 1e9
    traffic = measured_traffic(res["path"])
    kernel = KERNELS[res["path"]]
    out = {
        "metric": "Mcell-updates/s",
        "value": round(mcells, 1),
        "unit": "Mcell/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "256^3 FP32 Yee+CPML npml=12, 40 random dielectric boxes, 32x32 Jx "
                        "sheet, 8-frequency Huygens-box DFT [32,223]^3",
            "grid": list(wl.shape),
            "npml": wl.npml,
            "n2f_angles": "181 theta x 2 phi, R=1 m",
            "step_path": res["path"],
        },
        "roofline": {
            "bound": "hbm",
            "kernel": kernel,
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "bytes_alg_per_step": b_alg,
            "sweep_us_per_step": round(step_kern_s * 1e6, 2),
            "launches_per_step": round(res["launches"] / args.steps, 3),
            "timed_launches": res["timed"],
            # achieved/frac count algorithmic bytes; these count the bytes
            # the counters saw (traffic) over the same sweep time
            "traffic_gbs": None if traffic is None else round(traffic / step_kern_s / 1e9, 1),
            "traffic_frac": None if traffic is None
            else round(traffic / step_kern_s / 1e9 / HBM_PEAK_GBS, 4),
        },
        "n2f_ms": round(res["n2f_ms"], 2),
        "n2f_freqs": wl.n_freq,
        "farfield_finite": res["finite"],
        "device": res["device"],
    }
    if args.accuracy_steps > 0:
        out["fp64_check"] = run_accuracy(wl, res["tuned"], args.accuracy_steps,
                                         None if args.path == "auto" else args.path)
        print(f"[bench] FP32 vs FP64 after {args.accuracy_steps} steps: max peak-normalised "
              f"error {out['fp64_check']['max']:.2e}", file=sys.stderr, flush=True)
    if not args.no_cpu:
        cpu_mcells, cpu_s = run_cpu_baseline(wl, args.cpu_steps, args.cpu_warmup)
        out["cpu_baseline"] = {
            "value": round(cpu_mcells, 3),
            "unit": "Mcell/s",
            "cores": 1,
            "kind": "port",
            "sample": f"{args.cpu_steps} steps after {args.cpu_warmup} warm-up steps of the same "
                      f"256^3 workload (NumPy face-CPML restatement, {cpu_s:.1f} s), host has "
                      f"{os.cpu_count()} CPUs",
        }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
