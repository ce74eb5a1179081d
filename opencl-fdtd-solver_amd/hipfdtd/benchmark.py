"""Sustained-throughput benchmark of the HIP engine (benchmark CLI).

Restates the reference's GPU benchmark (``benchmarks/benchmark_vs_meep.py``:
``run_opencl_benchmark`` and ``main``, lines 62-135 and 244-331): a grid with
a Gaussian-modulated sine Ex sheet at z = Nz - npml - 2 (f = 6 GHz, 20 %
bandwidth, dl = 1 mm), no monitors, an explicit warm-up, then ``repeats``
timed windows of ``steps`` steps; it prints the median / min / max MCUPS and
writes a ``key=value`` results file.  The Meep comparison of the reference is
not part of this build (it needs Docker and conda-forge); ``--skip-meep`` is
accepted for command-line compatibility.

``python -m hipfdtd.benchmark --shape 256 --steps 100 --repeats 3``
"""

from __future__ import annotations

import argparse
import math
import statistics
import time
import warnings
from pathlib import Path

from . import _lib
from .engine import HIPFDTD

DEFAULT_SHAPE = (600, 600, 600)
DEFAULT_STEPS = 100
DEFAULT_WARMUP = 10
DEFAULT_REPEATS = 3
DEFAULT_NPML = 25


def parse_shape(text: str) -> tuple:
    parts = [int(p) for p in text.lower().replace("x", ",").split(",")]
    if len(parts) == 1:
        return parts[0], parts[0], parts[0]
    if len(parts) != 3:
        raise argparse.ArgumentTypeError("shape must be N or Nx,Ny,Nz")
    return tuple(parts)


def run_benchmark(shape=DEFAULT_SHAPE, steps=DEFAULT_STEPS, warmup=DEFAULT_WARMUP,
                  repeats=DEFAULT_REPEATS, npml=DEFAULT_NPML, fused=None) -> dict:  # fmt: skip
    """Timed windows of the HIP engine; returns the reference's result fields."""
    dl = 1e-3
    freq = 6e9
    fwidth = 0.2 * freq
    sim = HIPFDTD(shape, dl, npml=npml, fused=fused)
    needed = HIPFDTD.estimate_device_memory_bytes(shape, npml, fused=sim.fused) / 1024**3
    total_b, _, name = _lib.device_info(sim.device)
    budget = HIPFDTD.device_memory_budget_bytes(sim.device) / 1024**3
    print(f"Device:  {name}")
    print(f"VRAM:    {total_b / 1024**3:.2f} GB total, {budget:.2f} GB usable budget")
    print(f"Model:   ~{needed:.2f} GB float32 fields + face-local CPML buffers")
    z_src = shape[2] - npml - 2
    t0 = 5.0 / (math.pi * fwidth)
    sigma = 1.0 / (math.pi * fwidth)

    def source_cb(f):
        amp = math.exp(-0.5 * ((f.t - t0) / sigma) ** 2) * math.sin(2 * math.pi * freq * f.t)
        f.add_source_Ex(z_src, amp)

    sim.add_source(source_cb)
    cells = shape[0] * shape[1] * shape[2]
    times, rates = [], []
    with warnings.catch_warnings():
        warnings.simplefilter("ignore", DeprecationWarning)  # add_source_Ex, as the reference uses it
        sim.run(warmup)
        sim.synchronize()
        for rep in range(repeats):
            start = time.perf_counter()
            sim.run(steps)
            sim.synchronize()
            dur = time.perf_counter() - start
            times.append(dur)
            rates.append(cells * steps / dur / 1e6)
            print(f"  HIP window {rep + 1}/{repeats}: {dur:.3f}s ({rates[-1]:.1f} MCUPS)")
    sim.close()
    return {"device": name, "mem_gb": needed, "budget_gb": budget, "total_gb": total_b / 1024**3,
            "times": times, "mcups": rates, "time_s": statistics.median(times),
            "mcups_median": statistics.median(rates), "mcups_min": min(rates),
            "mcups_max": max(rates), "path": "fused" if sim.fused else "split"}  # fmt: skip


def main(argv=None) -> dict:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--shape", type=parse_shape, default=DEFAULT_SHAPE,
                    help="Grid size N or Nx,Ny,Nz (default 600)")  # fmt: skip
    ap.add_argument("--steps", type=int, default=DEFAULT_STEPS)
    ap.add_argument("--warmup", type=int, default=DEFAULT_WARMUP)
    ap.add_argument("--repeats", type=int, default=DEFAULT_REPEATS)
    ap.add_argument("--npml", type=int, default=DEFAULT_NPML)
    ap.add_argument("--path", choices=["auto", "fused", "split"], default="auto")
    ap.add_argument("--skip-meep", action="store_true", help="accepted; Meep is not run")
    ap.add_argument("--out", default="last_hip_benchmark.txt", help="key=value results file")
    args = ap.parse_args(argv)
    shape = args.shape
    cells = shape[0] * shape[1] * shape[2]
    print("=" * 60)
    print("HIP FDTD SUSTAINED THROUGHPUT BENCHMARK")
    print(f"Grid size: {shape[0]}x{shape[1]}x{shape[2]} = {cells / 1e6:.2f}M cells")
    print(f"Steps/window: {args.steps}  warmup: {args.warmup}  repeats: {args.repeats}")
    print("=" * 60 + "\n")
    r = run_benchmark(shape, args.steps, args.warmup, args.repeats, args.npml,
                      fused={"auto": None, "fused": True, "split": False}[args.path])  # fmt: skip
    print(f"\n[OK] HIP sustained median: {r['time_s']:.2f}s ({r['mcups_median']:.1f} MCUPS; "
          f"range {r['mcups_min']:.1f}-{r['mcups_max']:.1f})")  # fmt: skip
    lines = [f"device={r['device']}", f"shape={shape[0]}x{shape[1]}x{shape[2]}",
             f"cells_m={cells / 1e6:.2f}", f"mem_gb={r['mem_gb']:.2f}",
             f"budget_gb={r['budget_gb']:.2f}", f"steps={args.steps}", f"warmup={args.warmup}",
             f"repeats={args.repeats}", f"path={r['path']}", f"hip_s_median={r['time_s']:.4f}",
             f"hip_mcups_median={r['mcups_median']:.4f}", f"hip_mcups_min={r['mcups_min']:.4f}",
             f"hip_mcups_max={r['mcups_max']:.4f}"]  # fmt: skip
    Path(args.out).write_text("\n".join(lines) + "\n", encoding="utf-8")
    print(f"Wrote results to {args.out}")
    return r


if __name__ == "__main__":
    main()
