"""Pair sweep (two Yee steps per launch): bit-exact against the march path.

Every cell the pair sweep stores goes through the same per-cell arithmetic
as a single step, so fields, psi (through the fields of later steps) and
every monitor's sums must equal the march path's bit for bit, on lossless
and lossy grids, with and without PML, with sources in both steps, odd step
counts (the last step single) and more sources than the sweep applies
itself (two single steps).  One case is also held against the oracle.
"""

from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from oracle import fdtd_oracle as fo

pytestmark = pytest.mark.gpu

FIELDS = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")


def hip():
    import hipfdtd

    return hipfdtd


def _media(shape, seed, lossy):
    rng = np.random.default_rng(seed)
    eps = rng.uniform(1.0, 6.0, size=shape).astype(np.float32)
    sig = rng.uniform(0.0, 2.0, size=shape).astype(np.float32) if lossy else None
    return eps, sig


def _sources(sim, shape, n):
    """n current sources (a rim-tapered Jx sheet first), all live every step."""
    nx, ny, nz = shape
    boxes = [("y", (nx // 3, nx // 3 + 1, ny // 2, ny // 2 + 1, 5, 6)),
             ("z", (nx - 4, nx - 3, 2, 3, nz - 3, nz - 2)),
             ("x", (1, nx - 1, ny // 4, ny // 4 + 1, nz // 3, nz // 3 + 1)),
             ("y", (2, 3, 0, ny, 0, nz))]  # fmt: skip

    def src(f):
        a = fo.gaussian_sine(f.t, 3e10, 1.5e10)
        f.add_source_Jx(nz // 2, a, 2, nx - 3, 3, ny - 2, rim_taper=True, rim_edge=0.7)
        for q in range(n - 1):
            comp, box = boxes[q % len(boxes)]
            f.add_source_J(comp, (0.5 - 0.2 * q) * a, box)

    sim.add_source(src)


def _pair_and_march(shape, npml, lossy, nsrc, steps, planes=0, monitors=None):
    eps, sig = _media(shape, 21, lossy)
    out = {}
    for path in ("march", "pair"):
        sim = hip().HIPFDTD(shape, 1e-3, npml=npml, path=path)
        assert sim.path == path
        if path == "pair":
            sim.set_pair_planes(planes)
        sim.set_epsilon(eps, sig)
        _sources(sim, shape, nsrc)
        mons = monitors(sim) if monitors else []
        sim.run(steps)
        out[path] = (sim, mons)
    return out


def _assert_fields_equal(a, b):
    fa, fb = a.fields(), b.fields()
    for n in FIELDS:
        np.testing.assert_array_equal(fa[n], fb[n], err_msg=n)
    assert a.t == b.t and a.step_num == b.step_num


@pytest.mark.parametrize("shape,npml,lossy,nsrc,steps,planes", [
    ((45, 35, 100), 6, False, 3, 13, 0),
    ((45, 35, 100), 6, True, 4, 12, 5),
    ((20, 18, 16), 0, False, 1, 9, 0),
    ((70, 130, 125), 10, True, 2, 8, 24),
    ((33, 64, 61), 8, False, 4, 11, 1),
    ((64, 64, 64), 12, False, 3, 10, 1000)])  # fmt: skip
def test_pair_matches_march_bitwise(shape, npml, lossy, nsrc, steps, planes):
    print(f"[pair] shape={shape} npml={npml} lossy={lossy} nsrc={nsrc} steps={steps} "
          f"planes={planes}", flush=True)  # fmt: skip
    r = _pair_and_march(shape, npml, lossy, nsrc, steps, planes)
    _assert_fields_equal(r["march"][0], r["pair"][0])


def test_pair_matches_oracle():
    shape, npml = (37, 29, 70), 5
    print(f"[pair-oracle] shape={shape} npml={npml}", flush=True)
    eps, sig = _media(shape, 11, True)
    sim = hip().HIPFDTD(shape, 1e-3, npml=npml, path="pair")
    ora = fo.OracleFDTD(shape, 1e-3, npml=npml)
    for s in (sim, ora):
        s.set_epsilon(eps, sig)
        _sources(s, shape, 3)
    sim.run(24)
    for _ in range(24):
        ora.step()
    a, b = sim.fields(), ora.fields()
    num = sum(float(np.sum((a[n].astype(np.float64) - b[n]) ** 2)) for n in FIELDS)
    den = sum(float(np.sum(b[n].astype(np.float64) ** 2)) for n in FIELDS)
    assert den > 0 and (num / den) ** 0.5 < 1e-5


def test_pair_monitors_match_march_bitwise():
    """Huygens box, plane DFT and probe read the state after every step:
    after the first step of a pair from the watched cells of the shadow
    slab."""
    from hipfdtd.dft import HIPPlaneDFTMonitor
    from hipfdtd.monitors import HIPNear2FarMonitor
    from hipfdtd.probes import HIPPointProbe

    shape, dl = (40, 52, 70), 1e-3

    def monitors(sim):
        return [HIPNear2FarMonitor(sim, [s / 2 * dl for s in shape], [(s - 20) * dl for s in shape],
                                   freqs=[2.5e10, 3e10]),
                HIPNear2FarMonitor(sim, [s / 2 * dl for s in shape], [(s - 24) * dl for s in shape],
                                   freq=3e10, buffer_steps=0),
                HIPPlaneDFTMonitor(sim, "y", 26, ("Ex", "Hz"), freqs=(3e10,)),
                HIPPointProbe(sim, (21, 17, 35), "Ez", freqs=(2.8e10, 3e10))]  # fmt: skip

    print(f"[pair-monitors] shape={shape}", flush=True)
    r = _pair_and_march(shape, 6, False, 2, 17, monitors=monitors)
    (a, ma), (b, mb) = r["march"], r["pair"]
    _assert_fields_equal(a, b)
    obs = np.array([[0.3, 0.4, 0.5], [1.0, 0.0, 0.0], [0.0, -2.0, 0.1]])
    for x, y in zip(ma[:2], mb[:2]):
        np.testing.assert_array_equal(x.get_farfields(obs), y.get_farfields(obs))
    for x, y in zip(ma[2:], mb[2:]):
        np.testing.assert_array_equal(x.dft(), y.dft())


def test_pair_more_sources_than_sweep_takes_single_steps():
    """Five sources a step exceed the four the sweep applies: the pair path
    runs the two steps singly, with the same result."""
    shape = (32, 30, 40)
    print(f"[pair-5src] shape={shape}", flush=True)
    r = _pair_and_march(shape, 4, False, 5, 8)
    _assert_fields_equal(r["march"][0], r["pair"][0])


def test_pair_callbacks_that_read_fields_raise():
    sim = hip().HIPFDTD((16, 16, 16), 1e-3, npml=2, path="pair")
    sim.add_source(lambda f: f.add_source_J("x", float(f.Ex[8, 8, 8]), (8, 9, 8, 9, 8, 9)))
    with pytest.raises(RuntimeError, match="pair"):
        sim.run(2)


def test_pair_path_falls_back_for_fp64():
    sim = hip().HIPFDTD((16, 16, 16), 1e-3, npml=2, dtype=np.float64, path="pair")
    assert sim.path == "march"


def test_pair_entry_points_reject_bad_arguments():
    _lib = hip()._lib
    sim = hip().HIPFDTD((16, 16, 16), 1e-3, npml=2, path="pair")
    lib, h = sim._lib, sim._handle
    assert lib.hfdtd_set_view(h, 1) == _lib.EINVAL  # no watched boxes yet
    boxes = (C.c_int * 63)(*([0, 16, 0, 16, 0, 16, 0] * 9))
    assert lib.hfdtd_set_watch(h, boxes, 9) == _lib.EINVAL
    assert lib.hfdtd_set_pair_planes(h, -1) == _lib.EINVAL
    assert lib.hfdtd_set_pair_planes(h, 1025) == _lib.EINVAL
    src = (_lib.Source * 5)()
    for s in src:
        s.comp, s.mode, s.i0, s.i1, s.j0, s.j1, s.k0, s.k1 = 0, 1, 4, 5, 4, 5, 4, 5
    assert lib.hfdtd_step_pair(h, src, 5, src, 0) == _lib.EUNSUPPORTED
    assert lib.hfdtd_step_pair(h, src, 4, src, 4) == _lib.OK
    assert lib.hfdtd_set_watch(h, boxes, 1) == _lib.OK
    assert lib.hfdtd_set_view(h, 1) == _lib.OK
    assert lib.hfdtd_set_view(h, 0) == _lib.OK


def test_pair_full_size_matches_march():
    """256^3 with PML 12, the benchmark's grid: a few pairs against the march path."""
    shape = (256, 256, 256)
    print(f"[pair-full] shape={shape}", flush=True)
    r = _pair_and_march(shape, 12, False, 2, 6)
    _assert_fields_equal(r["march"][0], r["pair"][0])


@pytest.mark.parametrize("first", ["split", "march-read"])
def test_switch_to_pair_after_other_steps_matches_march(first):
    """Steps on the split path, or march steps whose callback reads a field,
    leave the H half of their step done; after set_path('pair') both steps'
    sources must still reach the pair sweep.  Both engines run the same
    first steps, then one goes on with the march sweep and one with the
    pair sweep; bitwise equal."""
    shape, npml = (40, 36, 72), 6
    print(f"[pair-switch] first={first} shape={shape}", flush=True)
    eps, _ = _media(shape, 5, False)
    out = []
    for then in ("march", "pair"):
        sim = hip().HIPFDTD(shape, 1e-3, npml=npml, path="march")
        sim.set_epsilon(eps)

        def src(f):
            a = fo.gaussian_sine(f.t, 3e10, 1.5e10)
            if first == "march-read" and f.step_num < 3:
                f.read_point("Ex", 20, 18, 36)  # runs the H half before the sources
            f.add_source_J("y", a, (20, 21, 18, 19, 30, 31))
            f.add_source_Jx(36, 0.5 * a, 10, 30, 9, 27, rim_taper=True)

        sim.add_source(src)
        sim.set_path("split" if first == "split" else "march")
        sim.run(3)
        sim.set_path(then)
        assert sim.path == then
        sim.run(10)
        out.append(sim)
    _assert_fields_equal(out[0], out[1])


def test_pair_callbacks_see_step_num():
    """A source keyed on step_num gets the same inputs on the pair path as on
    the march path (the second step of a pair sees step_num n + 1)."""
    shape = (32, 30, 40)
    print(f"[pair-step-num] shape={shape}", flush=True)
    res, seen = {}, {}
    for path in ("march", "pair"):
        sim = hip().HIPFDTD(shape, 1e-3, npml=4, path=path)
        log = []

        def src(f, log=log):
            log.append((f.step_num, f.t))
            f.add_source_J("z", 1.0 + (f.step_num % 3), (16, 17, 15, 16, 20, 21))

        sim.add_source(src)
        sim.run(9)
        res[path], seen[path] = sim, log
    assert seen["pair"] == seen["march"]
    assert [n for n, _ in seen["march"]] == list(range(9))
    _assert_fields_equal(res["march"], res["pair"])
