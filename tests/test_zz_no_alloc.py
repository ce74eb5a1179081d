"""The allocation-free far field (``hfdtd_n2f_farfield_noalloc``) against
``hfdtd_n2f_farfield``, called through the C ABI.

Both run the same kernels on the same face sums, so every call must give
the same E/H bit for bit: two calls in a row on a large monitor, then on a
smaller one (fewer face samples and points), so the second call runs on
scratch the first one left filled.  Requests beyond the scratch size are
refused with an error instead of running.
"""

from __future__ import annotations

import ctypes as C
import math

import numpy as np
import pytest

from oracle import fdtd_oracle as fo

pytestmark = pytest.mark.gpu

SCRATCH_DOUBLES = 1 << 20


def hip():
    import hipfdtd

    return hipfdtd


def _points(n, R, seed):
    rng = np.random.default_rng(seed)
    v = rng.normal(size=(n, 3))
    return np.ascontiguousarray(R * v / np.linalg.norm(v, axis=1, keepdims=True), np.float64)


def _call(mon, fn, freq, pts, l_sign=-1):
    lib = mon._lib
    k = float(np.float32(mon.omegas[freq] / fo.C0))
    eta = float(np.float32(fo.ETA0))
    eh = np.full(len(pts) * 12, np.nan)
    rc = getattr(lib, fn)(mon._handle, freq, pts.ctypes.data_as(C.c_void_p), len(pts), k, eta,
                          l_sign, 0, eh.ctypes.data_as(C.c_void_p))
    return rc, eh


def test_noalloc_farfield_matches_allocating_entry_point():
    hipfdtd = hip()
    shape, dl = (64, 60, 72), 1e-3
    sim = hipfdtd.HIPFDTD(shape, dl, npml=8)
    eps = np.ones(shape, np.float32)
    eps[24:40, 20:40, 30:44] = 3.0
    sim.set_epsilon(eps)
    sim.add_source(lambda f: f.add_source_Jx(36, fo.gaussian_sine(f.t, 3e10, 1.5e10), 28, 36,
                                             26, 34, rim_taper=True))  # fmt: skip
    c = [s * dl / 2 for s in shape]
    big = hipfdtd.HIPNear2FarMonitor(sim, c, [(s - 20) * dl for s in shape], freqs=[2.5e10, 3e10])
    small = hipfdtd.HIPNear2FarMonitor(sim, c, [12 * dl] * 3, freq=3e10)
    sim.run(120)
    print(f"[noalloc] n_face big={big.n_face_samples} small={small.n_face_samples}", flush=True)
    calls = [(big, 1, _points(300, 2.0, 1), -1), (small, 0, _points(40, 1.0, 2), +1),
             (big, 0, _points(17, 3.0, 3), +1)]  # fmt: skip
    for mon, freq, pts, sign in calls:
        rc_a, a = _call(mon, "hfdtd_n2f_farfield", freq, pts, sign)
        rc_b, b = _call(mon, "hfdtd_n2f_farfield_noalloc", freq, pts, sign)
        assert rc_a == 0 and rc_b == 0, hipfdtd._lib.last_error()
        assert np.all(np.isfinite(b)) and np.max(np.abs(b)) > 0
        np.testing.assert_array_equal(a, b)


def test_noalloc_farfield_refuses_requests_beyond_its_scratch():
    hipfdtd = hip()
    _lib = hipfdtd._lib
    shape, dl = (40, 40, 40), 1e-3
    sim = hipfdtd.HIPFDTD(shape, dl, npml=4)
    mon = hipfdtd.HIPNear2FarMonitor(sim, [20 * dl] * 3, [30 * dl] * 3, freq=3e10)
    nblk = math.ceil(mon.n_face_samples / 2048)
    fits = SCRATCH_DOUBLES // (12 * nblk + 15)
    rc, _ = _call(mon, "hfdtd_n2f_farfield_noalloc", 0, _points(fits + 1, 1.0, 5))
    assert rc == _lib.EINVAL and "scratch" in _lib.last_error()
    rc, eh = _call(mon, "hfdtd_n2f_farfield_noalloc", 0, _points(fits, 1.0, 5))
    assert rc == 0 and np.all(np.isfinite(eh))
