"""CPML 1-D profiles computed by the library's host code (no GPU needed).

Same signatures and meaning as the reference's cpml.py:68-150; the numbers come
from ``hfdtd_cpml_profile`` in ``libhipfdtd.so``, the same routine the engine
uses to fill its device profiles.
"""

from __future__ import annotations

import ctypes as C
from typing import NamedTuple

import numpy as np

from . import _lib
from .constants import ETA0

CPML_M = 3
CPML_KAPPA_MAX = 15.0
CPML_ALPHA_MAX_OVER_ETA0 = 0.05


class CpmlAxisProfile(NamedTuple):
    b: np.ndarray
    c: np.ndarray
    kappa: np.ndarray


class CpmlProfiles(NamedTuple):
    e: tuple
    h: tuple


def sigma_max(dl: float, *, m: int = CPML_M) -> float:
    """σ_max = 0.8(m+1)/(η0 Δ)."""
    return 0.8 * (m + 1) / (ETA0 * float(dl))


def build_cpml_1d(n, *, npml, dl, dt, node_offset, dtype=np.float32) -> CpmlAxisProfile:
    dt_ = np.dtype(dtype)
    if dt_ not in (np.dtype(np.float32), np.dtype(np.float64)):
        raise ValueError("CPML profiles are float32 or float64")
    b = np.empty(int(n), dtype=dt_)
    c = np.empty(int(n), dtype=dt_)
    k = np.empty(int(n), dtype=dt_)
    rc = _lib.load().hfdtd_cpml_profile(
        int(n), int(npml), float(dl), float(dt), float(node_offset),
        int(dt_ == np.float64), b.ctypes.data_as(C.c_void_p),
        c.ctypes.data_as(C.c_void_p), k.ctypes.data_as(C.c_void_p),
    )  # fmt: skip
    _lib.check(rc, "hfdtd_cpml_profile")
    return CpmlAxisProfile(b, c, k)


def build_cpml_profiles(shape, *, npml, dl, dt, dtype=np.float32) -> CpmlProfiles:
    """E-node (offset 0) and H-node (offset ½) profiles for x, y, z."""
    kw = dict(npml=npml, dl=dl, dt=dt, dtype=dtype)
    e = tuple(build_cpml_1d(int(n), node_offset=0.0, **kw) for n in shape)
    h = tuple(build_cpml_1d(int(n), node_offset=0.5, **kw) for n in shape)
    return CpmlProfiles(e=e, h=h)
