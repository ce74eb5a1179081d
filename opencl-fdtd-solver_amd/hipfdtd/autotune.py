"""Autotuner for the fused sweep's tile shape and the batched sweep's schedule.

The reference has no tuner (its OpenCL kernels take the driver's work-group
choice).  Here the launch shape is a real trade-off: rows of j per block set
the LDS footprint and so the resident blocks per CU; planes of i per block set
how many blocks a grid yields and how often the two halo planes are re-read;
a persistent batched launch also picks its grid size.  The best choice
depends on the grid extents, the dtype, lossy or not, and the device.

``tune_engine(sim)`` times the split sweeps and every fused-sweep and
march-sweep candidate (HIP events on the engine stream) on a scratch engine
of the same shape and applies the fastest; ``tune_batch(batch)`` does the same for a
``BatchedFDTD`` (fused tiles, flat split sweeps or march tiles) before its
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

FP32_ROWS = (2, 4, 6, 8, 10, 12)
FP64_ROWS = (1, 2, 4)
PLANES = (-1, 8, 16, 24, 32, 48, 64)
MARCH_ROWS_FP32 = (1, 2)
MARCH_ROWS_FP64 = (1,)
MARCH_PLANES = (-1, 16, 24, 32)
# march tile widths (columns of k): 32-column tiles (two rows of j per
# wavefront slot, one row per wavefront) are candidates on FP32 grids
MARCH_WIDTHS_FP32 = (64, 32)
MARCH_WIDTHS_FP64 = (64,)
PAIR_PLANES = (16, 24, 32)
DEFAULT_PATHS = ("fused", "split", "march")
BATCH_ROWS_FP32 = (2, 4, 6, 8)
BATCH_ROWS_FP64 = (2, 4)
BATCH_PLANES = (8, 16, 32, 64)
# batched march candidates: (rows per wavefront, planes per chunk, tile width)
BATCH_MARCH_FP32 = ((1, 16, 64), (1, 24, 64), (1, 32, 64), (2, 32, 64), (1, 16, 32), (1, 24, 32))
BATCH_MARCH_FP64 = ((1, 16, 64), (1, 32, 64), (1, 48, 64))
# batched split sweeps: runs of 256 cells per work item
BATCH_SPLIT_RUNS = (1, 2, 4)


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
    """µs of sweep kernels per step (the split path launches two per step,
    the pair path one per two steps)."""
    sim.run(warmup)
    sim.synchronize()
    sim.kernel_timing(True)
    sim.run(steps)
    ms, _ = sim.kernel_timing_read()
    sim.kernel_timing(False)
    return ms / max(steps, 1) * 1e3


def apply(sim, result) -> None:
    """Apply a ``tune_engine`` result (path and launch shape) to another
    engine of the same shape and medium kind."""
    _apply(sim, result["path"], result["rows"], result["planes"], result.get("width") or 64)


def _apply(sim, path, r, p, w=64):
    sim.set_path(path)
    if path == "fused":
        sim.set_fused_tiling(r, p, 0)
    elif path == "march":
        sim.set_march_width(w)
        sim.set_march_tiling(r, p, 0, -2)
    elif path == "pair":
        sim.set_pair_planes(p)


def tune_engine(sim, *, rows=None, planes=None, march_rows=None, march_planes=None,
                march_widths=None, pair_planes=None, paths=DEFAULT_PATHS, steps=20, warmup=3,
                use_cache=True, write_cache=True, cache_file=None) -> dict:  # fmt: skip
    """Pick the step path and its launch shape for ``sim``, and apply them.

    Times, on a scratch engine of the same shape, the split sweeps (one
    candidate), the LDS-DMA fused sweep for ``rows`` (j rows per block) x
    ``planes`` candidates and the march sweep for ``march_rows`` (j rows per
    wavefront) x ``march_planes`` candidates on 64-column tiles, and for
    ``march_planes`` on the narrower ``march_widths`` (one row per
    wavefront, when 1 is among ``march_rows``); with "pair" in ``paths``
    (opt-in) also the two-steps-per-launch sweep for ``pair_planes``.  The
    timing runs without monitors: the pair sweep's record of watched cells
    for monitors is not included.  Returns {"path": "split" | "fused" |
    "march" | "pair", "rows", "planes", "width" (march tile width, else
    None), "us": fastest µs of sweep kernels per step, "table": [[path,
    rows, planes, width, µs], ...], "cached": bool};
    ``planes`` -1 is the grid-derived default.  One-launch candidates are
    skipped for grids beyond their slab limits.
    """
    from .engine import HIPFDTD

    shape = (sim.Nx, sim.Ny, sim.Nz)
    # a grid with a material index runs other march kernels: tune those
    indexed = not sim.lossy and sim.material_count() > 0
    kind = "engine" if tuple(paths) == DEFAULT_PATHS else "engine:" + "+".join(sorted(paths))
    if indexed:
        kind += ":indexed"
    key = cache_key(kind, [shape], sim.dtype, sim.npml, sim.lossy, sim.device_name)
    if use_cache:
        hit = load_cache(cache_file).get(key)
        if hit and "width" in hit:
            _apply(sim, hit["path"], hit["rows"], hit["planes"], hit["width"] or 64)
            return dict(hit, cached=True)
    fp64 = sim.dtype == np.dtype(np.float64)
    rows = rows or (FP64_ROWS if fp64 else FP32_ROWS)
    planes = planes or PLANES
    march_rows = march_rows or (MARCH_ROWS_FP64 if fp64 else MARCH_ROWS_FP32)
    march_planes = march_planes or MARCH_PLANES
    march_widths = march_widths or (MARCH_WIDTHS_FP64 if fp64 else MARCH_WIDTHS_FP32)
    pair_planes = pair_planes or PAIR_PLANES
    scratch = HIPFDTD(shape, sim.dl, npml=sim.npml, dtype=sim.dtype, device=sim.device,
                      fused=False)  # fmt: skip
    table = []
    try:
        if sim.lossy:
            scratch.set_epsilon(np.ones(shape, np.float32), np.zeros(shape, np.float32))
        elif indexed:
            scratch.set_epsilon(np.ones(shape, np.float32))  # one material: indexed
        def timed(path, r, p, w):
            return [path, r, p, w, round(_time_engine(scratch, steps, warmup), 3)]

        if "split" in paths:
            table.append(timed("split", None, None, None))
        scratch.set_path("fused")
        if "fused" in paths and scratch.fused:
            for r in rows:
                for p in planes:
                    scratch.set_fused_tiling(r, p, 0)
                    table.append(timed("fused", r, p, None))
        scratch.set_path("march")
        if "march" in paths and scratch.path == "march":
            for w in march_widths:
                scratch.set_march_width(w)
                for r in march_rows if w == 64 else [r for r in march_rows if r == 1]:
                    for p in march_planes:
                        scratch.set_march_tiling(r, p, 0, -2)
                        table.append(timed("march", r, p, w))
            scratch.set_march_width(64)
        scratch.set_path("pair")
        if "pair" in paths and scratch.path == "pair":
            for p in pair_planes:
                scratch.set_pair_planes(p)
                table.append(timed("pair", None, p, None))
    finally:
        scratch.close()
    path, r, p, w, us = min(table, key=lambda e: e[4])
    _apply(sim, path, r, p, w or 64)
    entry = {"path": path, "rows": r, "planes": p, "width": w, "us": us, "table": table,
             "when": time.strftime("%Y-%m-%dT%H:%M:%S")}  # fmt: skip
    if write_cache:
        save_cache({key: entry}, cache_file)
    return dict(entry, cached=False)


def tune_batch(batch, *, rows=None, planes=None, blocks=(0,), paths=("fused", "split", "march"),
               march=None, split_runs=None, steps=10, warmup=2, use_cache=True,
               write_cache=True, cache_file=None) -> dict:  # fmt: skip
    """Pick the path ('fused' tiles, 'split' flat sweeps or 'march' tiles)
    and its launch shape for a ``BatchedFDTD`` and apply them: (blocks,
    rows_j, planes_i) for the fused path, runs of 256 cells per work item
    from ``split_runs`` for the split path (in the table's planes column),
    (rows per wavefront, planes per chunk, tile width) from ``march`` for the
    march path.

    Must run before the batch's source / probe tables are set, while every
    simulation is still at step 0 with zero fields (stepping them then leaves
    them unchanged).  Returns the choice, its ms per step and the table
    [[path, blocks, rows, planes, width, ms], ...]; rows and planes do not
    apply to the split path, width only to the march path.
    """
    sims = batch.sims
    if batch._n_table or any(s.step_num for s in sims):
        raise ValueError("tune_batch runs on a fresh batch (before set_sources / set_probes)")
    s0 = sims[0]
    key = cache_key("batch", [(s.Nx, s.Ny, s.Nz) for s in sims], s0.dtype, s0.npml,
                    any(s.lossy for s in sims), s0.device_name)  # fmt: skip
    fp64 = s0.dtype == np.dtype(np.float64)

    def apply(path, b, r, p, w):
        batch.set_path(path)
        if path == "march":
            batch.set_march(r, p, 0, w)
        elif path == "split":
            batch.set_launch(b, 0, 0)
            batch.set_split(p or 1)
        else:
            batch.set_launch(b, r or 0, p or 0)

    if use_cache:
        hit = load_cache(cache_file).get(key)
        if hit and hit["path"] in paths and "width" in hit:
            apply(hit["path"], hit["blocks"], hit["rows"], hit["planes"], hit["width"])
            return dict(hit, cached=True)
    rows = rows or (BATCH_ROWS_FP64 if fp64 else BATCH_ROWS_FP32)
    planes = planes or BATCH_PLANES
    march = march or (BATCH_MARCH_FP64 if fp64 else BATCH_MARCH_FP32)
    split_runs = split_runs or BATCH_SPLIT_RUNS
    lib = batch._lib

    def timed(path, b, r, p, w):
        # per step = (one run of 2 n steps) - (one run of n steps), over n:
        # what a run costs once (the march path's padded copies, the set-up
        # of the work list) drops out, as it does over a long run
        apply(path, b, r, p, w)
        _lib.check(lib.hfdtd_batch_run(batch._handle, warmup, 0), "batch tune")
        t0 = time.perf_counter()
        _lib.check(lib.hfdtd_batch_run(batch._handle, steps, 0), "batch tune")
        t1 = time.perf_counter()
        _lib.check(lib.hfdtd_batch_run(batch._handle, 2 * steps, 0), "batch tune")
        t2 = time.perf_counter()
        per = ((t2 - t1) - (t1 - t0)) / steps
        if per <= 0.0:  # timing noise on very short runs
            per = (t2 - t1) / (2 * steps)
        return [path, b, r, p, w, round(per * 1e3, 4)]

    table = []
    for b in blocks:
        if "fused" in paths:
            table += [timed("fused", b, r, p, None) for r in rows for p in planes]
        if "split" in paths:
            table += [timed("split", b, None, r, None) for r in split_runs]
    if "march" in paths:
        table += [timed("march", 0, r, p, w) for r, p, w in march]
    path, b, r, p, w, ms = min(table, key=lambda e: e[5])
    apply(path, b, r, p, w)
    entry = {"path": path, "blocks": b, "rows": r, "planes": p, "width": w, "ms": ms,
             "table": table, "when": time.strftime("%Y-%m-%dT%H:%M:%S")}  # fmt: skip
    if write_cache:
        save_cache({key: entry}, cache_file)
    return dict(entry, cached=False)
