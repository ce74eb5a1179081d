"""Autotuner for the fused sweep's tile shape and the batched sweep's schedule.

The reference has no tuner (its OpenCL kernels take the driver's work-group
choice).  Here the launch shape is a real trade-off: rows of j per block set
the LDS footprint and so the resident blocks per CU; planes of i per block set
how many blocks a grid yields and how often the two halo planes are re-read;
a persistent batched launch also picks its grid size.  The best choice
depends on the grid extents, the dtype, lossy or not, and the device.

``tune_engine(sim)`` times the fused sweep (HIP events on the engine stream)
for every candidate on a scratch engine of the same shape and applies the
fastest; ``tune_batch(batch)`` does the same for a ``BatchedFDTD`` before its
tables are set.  All stepping happens on zero fields with no sources, which
leaves the fields exactly zero, so tuning never changes a simulation.
Results are cached in a JSON file keyed by (kind, shapes, dtype, npml, lossy,
device, library ABI); ``HIPFDTD_TUNE_CACHE`` names the file (default
``~/.cache/hipfdtd/tuning.json``; an unwritable location only disables the
cache).
"""

from __future__ import annotations

import hashlib
import json
import os
import tempfile
import time
from pathlib import Path

import numpy as np

from . import _lib

FP32_ROWS = (1, 2, 4, 6, 8)
FP64_ROWS = (1, 2, 4)
PLANES = (-1, 8, 16, 24, 32, 48, 64)
BATCH_ROWS_FP32 = (2, 4, 6, 8)
BATCH_ROWS_FP64 = (2, 4)
BATCH_PLANES = (8, 16, 32, 64)


def cache_path() -> Path:
    env = os.environ.get("HIPFDTD_TUNE_CACHE")
    if env:
        return Path(env)
    return Path.home() / ".cache" / "hipfdtd" / "tuning.json"


def load_cache(path=None) -> dict:
    p = Path(path) if path else cache_path()
    try:
        data = json.loads(p.read_text())
        return data if isinstance(data, dict) else {}
    except (OSError, ValueError):
        return {}


def save_cache(entries: dict, path=None) -> bool:
    """Merge ``entries`` into the cache file (atomic replace); False if the
    location is not writable."""
    p = Path(path) if path else cache_path()
    try:
        p.parent.mkdir(parents=True, exist_ok=True)
        data = load_cache(p)
        data.update(entries)
        fd, tmp = tempfile.mkstemp(dir=str(p.parent), prefix=".tuning.", suffix=".json")
        with os.fdopen(fd, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)
        os.replace(tmp, p)
        return True
    except OSError:
        return False


def cache_key(kind, shapes, dtype, npml, lossy, device_name) -> str:
    """Stable key; a batch's shapes (in order) are hashed."""
    shapes = [list(map(int, s)) for s in shapes]
    if len(shapes) == 1:
        sh = "x".join(map(str, shapes[0]))
    else:
        sh = "batch%d-%s" % (len(shapes),
                             hashlib.sha1(json.dumps(shapes).encode()).hexdigest()[:16])  # fmt: skip
    return "|".join([kind, sh, np.dtype(dtype).name, f"npml{int(npml)}",
                     "lossy" if lossy else "lossless", str(device_name),
                     f"abi{_lib.load().hfdtd_version()}"])  # fmt: skip


def _time_engine(sim, steps, warmup):
    for _ in range(warmup):
        sim.step()
    sim.synchronize()
    sim.kernel_timing(True)
    for _ in range(steps):
        sim.step()
    ms, n = sim.kernel_timing_read()
    sim.kernel_timing(False)
    return ms / max(n, 1) * 1e3


def tune_engine(sim, *, rows=None, planes=None, steps=20, warmup=3, use_cache=True,
                write_cache=True, cache_file=None) -> dict:  # fmt: skip
    """Pick (rows_j, planes_i) of the fused sweep for ``sim`` and apply it.

    Returns {"rows": r, "planes": p, "us": fastest µs per sweep, "table":
    [[r, p, µs], ...], "cached": bool}.  ``planes`` -1 is the grid-derived
    default.  Engines that run the split sweeps are left unchanged.
    """
    from .engine import HIPFDTD

    if not sim.fused:
        return {"rows": None, "planes": None, "us": None, "table": [], "cached": False}
    shape = (sim.Nx, sim.Ny, sim.Nz)
    key = cache_key("fused", [shape], sim.dtype, sim.npml, sim.lossy, sim.device_name)
    if use_cache:
        hit = load_cache(cache_file).get(key)
        if hit:
            sim.set_fused_tiling(hit["rows"], hit["planes"], 0)
            return dict(hit, cached=True)
    fp64 = sim.dtype == np.dtype(np.float64)
    rows = rows or (FP64_ROWS if fp64 else FP32_ROWS)
    planes = planes or PLANES
    scratch = HIPFDTD(shape, sim.dl, npml=sim.npml, dtype=sim.dtype, device=sim.device)
    try:
        if sim.lossy:
            scratch.set_epsilon(np.ones(shape, np.float32), np.zeros(shape, np.float32))
        table = []
        for r in rows:
            for p in planes:
                scratch.set_fused_tiling(r, p, 0)
                table.append([r, p, round(_time_engine(scratch, steps, warmup), 3)])
    finally:
        scratch.close()
    r, p, us = min(table, key=lambda e: e[2])
    sim.set_fused_tiling(r, p, 0)
    entry = {"rows": r, "planes": p, "us": us, "table": table,
             "when": time.strftime("%Y-%m-%dT%H:%M:%S")}  # fmt: skip
    if write_cache:
        save_cache({key: entry}, cache_file)
    return dict(entry, cached=False)


def tune_batch(batch, *, rows=None, planes=None, blocks=(0,), steps=10, warmup=2,
               use_cache=True, write_cache=True, cache_file=None) -> dict:  # fmt: skip
    """Pick (blocks, rows_j, planes_i) of a ``BatchedFDTD`` and apply it.

    Must run before the batch's source / probe tables are set, while every
    simulation is still at step 0 with zero fields (stepping them then leaves
    them unchanged).  Returns the choice, its ms per step and the table.
    """
    sims = batch.sims
    if batch._n_table or any(s.step_num for s in sims):
        raise ValueError("tune_batch runs on a fresh batch (before set_sources / set_probes)")
    s0 = sims[0]
    key = cache_key("batch", [(s.Nx, s.Ny, s.Nz) for s in sims], s0.dtype, s0.npml,
                    any(s.lossy for s in sims), s0.device_name)  # fmt: skip
    if use_cache:
        hit = load_cache(cache_file).get(key)
        if hit:
            batch.set_launch(hit["blocks"], hit["rows"], hit["planes"])
            return dict(hit, cached=True)
    fp64 = s0.dtype == np.dtype(np.float64)
    rows = rows or (BATCH_ROWS_FP64 if fp64 else BATCH_ROWS_FP32)
    planes = planes or BATCH_PLANES
    lib = batch._lib
    table = []
    for b in blocks:
        for r in rows:
            for p in planes:
                batch.set_launch(b, r, p)
                _lib.check(lib.hfdtd_batch_run(batch._handle, warmup, 0), "batch tune")
                t0 = time.perf_counter()
                _lib.check(lib.hfdtd_batch_run(batch._handle, steps, 0), "batch tune")
                table.append([b, r, p, round((time.perf_counter() - t0) / steps * 1e3, 4)])
    b, r, p, ms = min(table, key=lambda e: e[3])
    batch.set_launch(b, r, p)
    entry = {"blocks": b, "rows": r, "planes": p, "ms": ms, "table": table,
             "when": time.strftime("%Y-%m-%dT%H:%M:%S")}  # fmt: skip
    if write_cache:
        save_cache({key: entry}, cache_file)
    return dict(entry, cached=False)
