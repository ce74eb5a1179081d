"""GPU: the autotuner picks a launch shape without changing results.

* ``tune_engine`` on a fresh engine leaves its fields zero, applies a
  candidate (the split sweeps, a fused or a march tiling) from its table, is served
  from the cache the second time, and the tuned engine steps bit-identically
  to an untuned one on the same path (the tile shape only reorders
  independent cell updates);
* ``tune_batch`` on a fresh batch applies a schedule and refuses a batch
  whose tables are set.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import fdtd_oracle as fo

pytestmark = pytest.mark.gpu

SHAPE = (40, 48, 72)
DL = 1e-3


def src(f):
    f.add_source_J("z", fo.gaussian_sine(f.t, 3e10, 3e10), (20, 21, 24, 25, 36, 37))


def test_tune_engine_applies_and_caches(tmp_path):
    import hipfdtd
    from hipfdtd.autotune import tune_engine

    cache = tmp_path / "tune.json"
    sim = hipfdtd.HIPFDTD(SHAPE, DL, npml=6)
    res = tune_engine(sim, planes=(-1, 8, 24), march_planes=(-1, 4), steps=5, warmup=1,
                      cache_file=cache)  # fmt: skip
    print("[autotune] engine", res["path"], res["rows"], res["planes"], res["us"], flush=True)
    assert not res["cached"]
    assert [res["path"], res["rows"], res["planes"], res["width"], res["us"]] in res["table"]
    # split; fused 6 rows x 3 planes; march 2 rows x 2 planes on 64-column
    # tiles and 2 planes on 32-column tiles
    assert len(res["table"]) == 1 + 6 * 3 + 2 * 2 + 2
    assert {e[3] for e in res["table"] if e[0] == "march"} == {64, 32}
    assert sim.path == res["path"]
    assert cache.exists()
    again = tune_engine(sim, cache_file=cache)
    assert again["cached"] and again["path"] == res["path"] and again["planes"] == res["planes"]
    assert not np.any(sim.read_volume("Ex"))
    ref = hipfdtd.HIPFDTD(SHAPE, DL, npml=6, path=res["path"])
    for s in (sim, ref):
        s.add_source(src)
        s.run(25)
    for name in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz"):
        assert np.array_equal(sim.read_volume(name), ref.read_volume(name)), name


def test_tune_engine_with_the_pair_path(tmp_path):
    """The pair sweep is an opt-in candidate with a cache entry of its own;
    whichever path wins, the tuned engine steps like a plain march engine
    (the pair sweep equals two march steps bit for bit)."""
    import hipfdtd
    from hipfdtd.autotune import tune_engine

    cache = tmp_path / "tune.json"
    sim = hipfdtd.HIPFDTD(SHAPE, DL, npml=6)
    res = tune_engine(sim, march_rows=(2,), march_planes=(-1,), pair_planes=(8, 24),
                      paths=("march", "pair"), steps=6, warmup=2, cache_file=cache)  # fmt: skip
    print("[autotune] pair", res["path"], res["planes"], res["table"], flush=True)
    assert [e[0] for e in res["table"]] == ["march", "pair", "pair"]
    assert sim.path == res["path"] in ("march", "pair")
    plain = tune_engine(hipfdtd.HIPFDTD(SHAPE, DL, npml=6), march_planes=(-1,), steps=4,
                        warmup=1, cache_file=cache)  # fmt: skip
    assert not plain["cached"] and plain["path"] != "pair"
    assert not np.any(sim.read_volume("Ex"))
    ref = hipfdtd.HIPFDTD(SHAPE, DL, npml=6, path="march")
    for s in (sim, ref):
        s.add_source(src)
        s.run(26)
    for name in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz"):
        assert np.array_equal(sim.read_volume(name), ref.read_volume(name)), name


def test_tune_batch_applies_schedule(tmp_path):
    import hipfdtd
    from hipfdtd.autotune import tune_batch
    from hipfdtd.batch import BatchedFDTD

    sims = [hipfdtd.HIPFDTD(s, DL, npml=4) for s in ((24, 30, 40), (32, 20, 64), (20, 20, 20))]
    b = BatchedFDTD(sims)
    res = tune_batch(b, rows=(2, 4), planes=(8, 16), march=((1, 16, 32), (2, 8, 64)),
                     split_runs=(1, 4), steps=3, warmup=1, cache_file=tmp_path / "tune.json")  # fmt: skip
    print("[autotune] batch", res["path"], res["blocks"], res["rows"], res["planes"], res["ms"],
          flush=True)  # fmt: skip
    assert len(res["table"]) == 8
    fused = {(2, 8), (2, 16), (4, 8), (4, 16)}
    assert ((res["path"] == "split" and res["planes"] in (1, 4))
            or (res["path"] == "fused" and (res["rows"], res["planes"]) in fused)
            or (res["path"] == "march"
                and (res["rows"], res["planes"], res["width"]) in {(1, 16, 32), (2, 8, 64)}))  # fmt: skip
    assert b.path == res["path"]
    for s in sims:
        assert not np.any(s.read_volume("Ez"))
    b.set_sources([("z", (10, 11, 10, 11, 10, 11))] * 3, [lambda t: 1.0] * 3, 4)
    with pytest.raises(ValueError):
        tune_batch(b, use_cache=False, write_cache=False)
    b.run(4)
    assert all(np.any(s.read_volume("Ez")) for s in sims)
    b.close()


@pytest.mark.parametrize("path,per_step", [("split", 2), ("fused", 1), ("march", 1)])
def test_sampled_sweep_timing(path, per_step):
    """Timing every n-th sweep launch: the sample holds launches 0, n, 2n ...
    of all the sweep launches while timing is on (the split path's H and E
    sweeps alternate), each with a positive duration, and stepping is
    unchanged by it."""
    import hipfdtd

    sim = hipfdtd.HIPFDTD(SHAPE, DL, npml=6, path=path)
    ref = hipfdtd.HIPFDTD(SHAPE, DL, npml=6, path=path)
    for s in (sim, ref):
        s.add_source(src)
        s.run(3)
    steps, every = 10, 3
    sim.kernel_timing(True, every=every)
    sim.run(steps)
    ms, timed = sim.kernel_timing_read()
    launches = sim.kernel_timing_launches()
    sim.kernel_timing(False)
    print(f"[timing] {path}: {timed} of {launches} launches timed, {ms:.3f} ms", flush=True)
    assert launches == steps * per_step
    assert timed == -(-launches // every)
    assert ms > 0
    sim.run(2)  # timing off: no more launches counted or timed
    assert sim.kernel_timing_read()[1] == timed and sim.kernel_timing_launches() == launches
    ref.run(steps + 2)
    for name in ("Ex", "Hz"):
        assert np.array_equal(sim.read_volume(name), ref.read_volume(name)), name
    with pytest.raises(ValueError):
        sim.kernel_timing(True, every=0)
