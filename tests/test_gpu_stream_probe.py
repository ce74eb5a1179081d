"""The streaming-bandwidth probe (hfdtd_stream_probe) through the C ABI:
each kernel kind returns a positive launch time, and bad arguments are
rejected with EINVAL before anything is allocated or launched."""

from __future__ import annotations

import ctypes as C

import pytest

pytestmark = pytest.mark.gpu


def _lib():
    from hipfdtd import _lib

    return _lib


@pytest.mark.parametrize("kind", [0, 1, 2, 3])
def test_stream_probe_times_each_kind(kind):
    L = _lib()
    sec = C.c_double(-1.0)
    print(f"[stream-probe] kind={kind}", flush=True)
    L.check(L.load().hfdtd_stream_probe(0, kind, 1 << 20, 64, 2, C.byref(sec)), "probe")
    assert 0.0 < sec.value < 1.0


@pytest.mark.parametrize("kind,nbytes,blocks,reps", [
    (4, 1 << 20, 64, 1), (-1, 1 << 20, 64, 1), (0, 100, 64, 1), (0, (1 << 20) + 4, 64, 1),
    (0, 1 << 20, 0, 1), (0, 1 << 20, 64, 0)])  # fmt: skip
def test_stream_probe_rejects_bad_arguments(kind, nbytes, blocks, reps):
    L = _lib()
    sec = C.c_double(0.0)
    assert L.load().hfdtd_stream_probe(0, kind, nbytes, blocks, reps, C.byref(sec)) == L.EINVAL
    assert L.load().hfdtd_stream_probe(0, 0, 1 << 20, 64, 1, None) == L.EINVAL
