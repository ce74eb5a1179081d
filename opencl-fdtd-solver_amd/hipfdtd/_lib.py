"""ctypes binding of ``libhipfdtd.so`` (declared in ``include/hipfdtd.h``).

The library must exist: there is no CPU fallback.  ``load()`` raises
``RuntimeError`` naming the missing file, so a product path can never run on
anything but the HIP kernels.
"""

from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

LIB_PATH = Path(os.environ.get("HIPFDTD_LIB", Path(__file__).resolve().parent / "libhipfdtd.so"))

OK = 0
EINVAL = -1
ENOMEM = -2
EHIP = -3
EUNSUPPORTED = -4
ENODEV = -5

VOLUME_IDS = {
    "Ex": 0, "Ey": 1, "Ez": 2, "Hx": 3, "Hy": 4, "Hz": 5,
    "ce_x": 6, "ce_y": 7, "ce_z": 8, "ca_x": 9, "ca_y": 10, "ca_z": 11,
}  # fmt: skip

SRC_SOFT = 0
SRC_CURRENT = 1
ACC_REAL = 0
ACC_FP64 = 1
ACC_KAHAN = 2
PATH_SPLIT = 0
PATH_FUSED = 1
PATH_MARCH = 2
# single steps of the "pair" path (two steps per launch) take the march sweep
PATH_CODES = {"split": PATH_SPLIT, "fused": PATH_FUSED, "march": PATH_MARCH, "pair": PATH_MARCH}
PAIR_MAX_SOURCES = 4  # sources per step the pair sweep applies itself
PAIR_MAX_WATCH = 8    # watched boxes
DFT_DIRECT = 0
DFT_BUFFERED = 1


class Source(C.Structure):
    """``hfdtd_source``: one soft or current source on a box."""

    _fields_ = [
        ("comp", C.c_int),
        ("mode", C.c_int),
        ("i0", C.c_int), ("i1", C.c_int),
        ("j0", C.c_int), ("j1", C.c_int),
        ("k0", C.c_int), ("k1", C.c_int),
        ("taper", C.c_int),
        ("rim_edge", C.c_double),
        ("amp", C.c_double),
    ]  # fmt: skip


EXPORTS = {
    "hfdtd_last_error": (C.c_char_p, []),
    "hfdtd_version": (C.c_int, []),
    "hfdtd_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "hfdtd_device_info": (
        C.c_int,
        [C.c_int, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), C.c_char_p, C.c_int],
    ),
    "hfdtd_padded_begin": (C.c_int, [C.c_void_p]),
    "hfdtd_padded_end": (C.c_int, [C.c_void_p]),
    "hfdtd_padded_live": (C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    "hfdtd_stream_probe": (
        C.c_int, [C.c_int, C.c_int, C.c_uint64, C.c_int, C.c_int, C.POINTER(C.c_double)],
    ),
    "hfdtd_cpml_profile": (
        C.c_int,
        [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_int,
         C.c_void_p, C.c_void_p, C.c_void_p],
    ),  # fmt: skip
    "hfdtd_create": (
        C.c_int,
        [C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, C.c_int,
         C.POINTER(C.c_void_p)],
    ),  # fmt: skip
    "hfdtd_destroy": (C.c_int, [C.c_void_p]),
    "hfdtd_memory_bytes": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint64)]),
    "hfdtd_set_materials": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]),
    "hfdtd_write_volume": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p]),
    "hfdtd_read_volume": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p]),
    "hfdtd_read_point": (C.c_int, [C.c_void_p, C.c_int, C.c_int64, C.POINTER(C.c_double)]),
    "hfdtd_reset_fields": (C.c_int, [C.c_void_p]),
    "hfdtd_update_H": (C.c_int, [C.c_void_p]),
    "hfdtd_update_E": (C.c_int, [C.c_void_p]),
    "hfdtd_add_sources": (C.c_int, [C.c_void_p, C.POINTER(Source), C.c_int]),
    "hfdtd_step": (C.c_int, [C.c_void_p, C.POINTER(Source), C.c_int, C.c_int]),
    "hfdtd_synchronize": (C.c_int, [C.c_void_p]),
    "hfdtd_set_fused_tiling": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int]),
    "hfdtd_set_split_schedule": (C.c_int, [C.c_void_p, C.c_int]),
    "hfdtd_step_pair": (C.c_int, [C.c_void_p, C.POINTER(Source), C.c_int, C.POINTER(Source),
                                  C.c_int]),
    "hfdtd_set_watch": (C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.c_int]),
    "hfdtd_set_view": (C.c_int, [C.c_void_p, C.c_int]),
    "hfdtd_set_pair_planes": (C.c_int, [C.c_void_p, C.c_int]),
    "hfdtd_material_count": (C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    "hfdtd_ce_source": (C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    "hfdtd_set_march_tiling": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int]),
    "hfdtd_set_march_addressing": (C.c_int, [C.c_void_p, C.c_int]),
    "hfdtd_set_overlap": (C.c_int, [C.c_void_p, C.c_int]),
    "hfdtd_set_march_width": (C.c_int, [C.c_void_p, C.c_int]),
    "hfdtd_timing": (C.c_int, [C.c_void_p, C.c_int]),
    "hfdtd_timing_read": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int)]),
    "hfdtd_timing_launches": (C.c_int, [C.c_void_p, C.POINTER(C.c_longlong)]),
    "hfdtd_n2f_create": (
        C.c_int,
        [C.c_void_p, C.POINTER(C.c_int), C.c_int, C.c_int, C.POINTER(C.c_void_p)],
    ),
    "hfdtd_n2f_destroy": (C.c_int, [C.c_void_p]),
    "hfdtd_n2f_layout": (C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "hfdtd_n2f_accumulate": (C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    "hfdtd_n2f_snapshot": (C.c_int, [C.c_void_p]),
    "hfdtd_n2f_rel_change": (
        C.c_int,
        [C.c_void_p, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)],
    ),
    "hfdtd_n2f_read": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_void_p]),
    "hfdtd_n2f_farfield": (
        C.c_int,
        [C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_int, C.c_int,
         C.c_void_p],
    ),  # fmt: skip
    "hfdtd_n2f_farfield_noalloc": (
        C.c_int,
        [C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_int, C.c_int,
         C.c_void_p],
    ),  # fmt: skip
    "hfdtd_probe_create": (
        C.c_int,
        [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_void_p)],
    ),
    "hfdtd_probe_destroy": (C.c_int, [C.c_void_p]),
    "hfdtd_probe_accumulate": (C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    "hfdtd_probe_read": (C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    "hfdtd_n2f_set_buffering": (C.c_int, [C.c_void_p, C.c_int]),
    "hfdtd_n2f_flush": (C.c_int, [C.c_void_p]),
    "hfdtd_n2f_zrecord": (C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    "hfdtd_dft_plane_create": (
        C.c_int,
        [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
         C.POINTER(C.c_void_p)],
    ),  # fmt: skip
    "hfdtd_dft_destroy": (C.c_int, [C.c_void_p]),
    "hfdtd_dft_accumulate": (C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    "hfdtd_dft_flush": (C.c_int, [C.c_void_p]),
    "hfdtd_dft_reset": (C.c_int, [C.c_void_p]),
    "hfdtd_dft_shape": (
        C.c_int,
        [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
         C.POINTER(C.c_int)],
    ),  # fmt: skip
    "hfdtd_dft_read": (C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    "hfdtd_batch_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_int, C.POINTER(C.c_void_p)]),
    "hfdtd_batch_destroy": (C.c_int, [C.c_void_p]),
    "hfdtd_batch_set_sources": (
        C.c_int,
        [C.c_void_p, C.POINTER(Source), C.POINTER(C.c_double), C.c_int],
    ),
    "hfdtd_batch_set_probes": (
        C.c_int,
        [C.c_void_p, C.POINTER(C.c_int), C.c_int, C.c_int, C.POINTER(C.c_double), C.c_int],
    ),
    "hfdtd_batch_run": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    "hfdtd_batch_read_probes": (C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    "hfdtd_batch_set_launch": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int]),
    "hfdtd_batch_set_path": (C.c_int, [C.c_void_p, C.c_int]),
    "hfdtd_batch_set_march": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int]),
    "hfdtd_batch_set_march_padding": (C.c_int, [C.c_void_p, C.c_int]),
    "hfdtd_batch_set_split": (C.c_int, [C.c_void_p, C.c_int]),
    "hfdtd_adjoint_create": (
        C.c_int,
        [C.c_void_p, C.POINTER(C.c_int), C.c_int, C.c_int, C.POINTER(C.c_void_p)],
    ),
    "hfdtd_adjoint_destroy": (C.c_int, [C.c_void_p]),
    "hfdtd_adjoint_memory_bytes": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint64)]),
    "hfdtd_adjoint_grad_box": (C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    "hfdtd_adjoint_set_probes": (
        C.c_int,
        [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.c_int, C.POINTER(C.c_double)],
    ),
    "hfdtd_adjoint_probe_reset": (C.c_int, [C.c_void_p]),
    "hfdtd_adjoint_probe_read": (C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    "hfdtd_adjoint_set_objective_grad": (C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    "hfdtd_adjoint_save": (C.c_int, [C.c_void_p, C.c_int]),
    "hfdtd_adjoint_restore": (C.c_int, [C.c_void_p, C.c_int]),
    "hfdtd_adjoint_record": (C.c_int, [C.c_void_p, C.c_int]),
    "hfdtd_adjoint_forward": (
        C.c_int,
        [C.c_void_p, C.POINTER(Source), C.c_int, C.POINTER(C.c_double), C.c_int, C.c_int64,
         C.c_int, C.c_int, C.c_int],
    ),  # fmt: skip
    "hfdtd_adjoint_begin": (C.c_int, [C.c_void_p, C.c_int64]),
    "hfdtd_adjoint_reverse": (C.c_int, [C.c_void_p, C.c_int64, C.c_int64, C.c_int]),
    "hfdtd_adjoint_read_grad": (C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    "hfdtd_adjoint_read_volume": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p]),
}

_LIB = None


def load() -> C.CDLL:
    """Load the HIP library once; raise RuntimeError when it is missing."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not LIB_PATH.exists():
        raise RuntimeError(
            f"HIP library {LIB_PATH} is missing: build it with `python -m hipfdtd.build` "
            "(hipcc --offload-arch=gfx950). There is no CPU fallback."
        )
    lib = C.CDLL(str(LIB_PATH))
    for name, (res, args) in EXPORTS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = lib
    return lib


def last_error() -> str:
    msg = load().hfdtd_last_error()
    return msg.decode(errors="replace") if msg else ""


def check(rc: int, what: str = "") -> None:
    """Map a status code to the reference's exception types."""
    if rc == OK:
        return
    msg = last_error()
    text = f"{what}: {msg}" if what else msg
    if rc in (EINVAL, EUNSUPPORTED):
        raise ValueError(text)
    if rc == ENOMEM:
        raise MemoryError(text)
    raise RuntimeError(text)


def device_count() -> int:
    n = C.c_int(0)
    rc = load().hfdtd_device_count(C.byref(n))
    return int(n.value) if rc == OK else 0


def device_info(device: int = 0):
    """(total_bytes, free_bytes, name) of a HIP device."""
    tot = C.c_uint64(0)
    free = C.c_uint64(0)
    name = C.create_string_buffer(256)
    check(load().hfdtd_device_info(int(device), C.byref(tot), C.byref(free), name, 256),
          "hfdtd_device_info")  # fmt: skip
    return int(tot.value), int(free.value), name.value.decode(errors="replace")
