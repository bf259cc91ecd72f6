"""ctypes binding of ``librebop_gpu.so`` (the C ABI in include/rebop_gpu.h).

The library is built in-tree by ``rebop_amd.build``.  There is no CPU
fallback: if the library is missing or the device is not gfx950, the calls
raise.
"""

from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_char_p, c_double, c_int, c_int64, c_size_t, c_uint32, c_uint64, c_void_p
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "librebop_gpu.so"

REBOP_OK = 0
REBOP_ERR_SHAPE = 1
REBOP_ERR_INDEX = 2
REBOP_ERR_VALUE = 3
REBOP_ERR_DEVICE = 4
REBOP_ERR_ALLOC = 5
REBOP_ERR_RANGE = 6
REBOP_ERR_UNSUPPORTED = 7
REBOP_ERR_CAP = 8

REBOP_KERNEL_AUTO = 0
REBOP_KERNEL_PLAIN = 1
REBOP_KERNEL_PERSISTENT = 2
REBOP_KERNEL_LDS = 3
REBOP_KERNEL_SPECIALIZED = 16
REBOP_KERNEL_GROUP2 = 32
REBOP_KERNEL_GROUP4 = 64
REBOP_KERNEL_WHOLE = 128
REBOP_KERNEL_WIDE = 256
REBOP_KERNEL_COMPACT = 512  # reported: the LDS kernel ran with int16 counts

# rate-expression bytecode operations (token = op | arg << 8)
REBOP_EXPR_CONST = 1
REBOP_EXPR_SPECIES = 2
REBOP_EXPR_NEG = 3
REBOP_EXPR_ADD = 4
REBOP_EXPR_SUB = 5
REBOP_EXPR_MUL = 6
REBOP_EXPR_DIV = 7
REBOP_EXPR_POW = 8
REBOP_EXPR_MAX = 9
REBOP_EXPR_MIN = 10
REBOP_EXPR_EXP = 11

REBOP_SYSTEM_MAX_COUNT = 64  # largest define_system stoichiometric coefficient

# every entry point of include/rebop_gpu.h: name -> (restype, argtypes)
SIGNATURES: dict[str, tuple] = {
    "rebop_gillespie_new": (c_int, [POINTER(c_int64), c_uint32, c_int, POINTER(c_void_p)]),
    "rebop_gillespie_new_with_seed": (
        c_int,
        [POINTER(c_int64), c_uint32, c_int, c_uint64, POINTER(c_void_p)],
    ),
    "rebop_gillespie_free": (None, [c_void_p]),
    "rebop_gillespie_seed": (c_int, [c_void_p, c_uint64]),
    "rebop_gillespie_set_draw_cap": (c_int, [c_void_p, c_uint64]),
    "rebop_gillespie_nb_species": (c_uint32, [c_void_p]),
    "rebop_gillespie_nb_reactions": (c_uint32, [c_void_p]),
    "rebop_gillespie_add_reaction_lma": (
        c_int,
        [c_void_p, c_double, POINTER(c_uint32), c_uint32, POINTER(c_int64), c_uint32],
    ),
    "rebop_gillespie_add_reaction_lma_sparse": (
        c_int,
        [c_void_p, c_double, POINTER(c_uint32), POINTER(c_uint32), c_uint32, POINTER(c_int64), c_uint32],
    ),
    "rebop_gillespie_add_reaction_expr": (
        c_int,
        [c_void_p, POINTER(c_uint32), c_uint32, POINTER(c_double), c_uint32, POINTER(c_int64), c_uint32],
    ),
    "rebop_gillespie_prepare": (c_int, [c_void_p, c_int]),
    "rebop_gillespie_set_species": (c_int, [c_void_p, POINTER(c_int64), c_uint32]),
    "rebop_gillespie_get_species": (c_int, [c_void_p, POINTER(c_int64), c_uint32]),
    "rebop_run": (
        c_int,
        [c_void_p, c_double, c_uint64, c_uint64, c_uint64, POINTER(c_int64), POINTER(c_int64),
         POINTER(c_double), c_void_p],
    ),
    "rebop_run_device": (
        c_int,
        [c_void_p, c_double, c_uint64, c_uint64, c_uint64, c_void_p, c_void_p, c_void_p, c_int,
         c_void_p, c_void_p],
    ),
    "rebop_run_device_noalloc": (
        c_int,
        [c_void_p, c_double, c_uint64, c_uint64, c_uint64, c_void_p, c_void_p, c_void_p, c_int,
         c_void_p, c_void_p],
    ),
    "rebop_run_all_events": (
        c_int,
        [c_void_p, c_double, c_uint64, c_uint64, POINTER(c_uint64), POINTER(c_double), POINTER(c_int64), c_void_p],
    ),
    "rebop_run_all_events_device": (
        c_int,
        [c_void_p, c_double, c_uint64, c_uint64, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p],
    ),
    "rebop_autotune": (c_int, [c_void_p, c_double, c_uint64, c_uint64, c_void_p]),
    "rebop_run_sweep_device": (
        c_int,
        [c_void_p, c_uint64, c_uint64, c_void_p, c_void_p, c_double, c_uint64, c_uint64, c_void_p,
         c_void_p, c_int, c_void_p, c_void_p],
    ),
    "rebop_run_jobs_device": (c_int, [c_void_p, c_uint32, c_void_p, c_void_p, POINTER(c_uint64)]),
    "rebop_batch_new": (c_int, [c_void_p, c_uint64, c_uint64, POINTER(c_int64), POINTER(c_void_p)]),
    "rebop_batch_free": (None, [c_void_p]),
    "rebop_batch_advance_until": (c_int, [c_void_p, c_double, c_void_p]),
    "rebop_batch_get_species": (c_int, [c_void_p, POINTER(c_int64)]),
    "rebop_batch_set_species": (c_int, [c_void_p, POINTER(c_int64)]),
    "rebop_batch_get_time": (c_int, [c_void_p, POINTER(c_double)]),
    "rebop_batch_set_time": (c_int, [c_void_p, POINTER(c_double)]),
    "rebop_batch_get_draws": (c_int, [c_void_p, POINTER(c_uint64)]),
    "rebop_batch_set_draws": (c_int, [c_void_p, POINTER(c_uint64)]),
    "rebop_system_new": (c_int, [POINTER(c_int64), c_uint32, c_uint64, POINTER(c_void_p)]),
    "rebop_system_add_reaction": (
        c_int,
        [c_void_p, c_double, POINTER(c_uint32), POINTER(c_uint32), c_uint32,
         POINTER(c_uint32), POINTER(c_uint32), c_uint32],
    ),
    "rebop_last_error": (c_char_p, []),
    "rebop_device_info": (c_int, [c_char_p, c_size_t, POINTER(c_int)]),
}


class Stats(ctypes.Structure):
    """``rebop_stats`` of include/rebop_gpu.h."""

    _fields_ = [
        ("events", c_uint64),
        ("draws", c_uint64),
        ("trajectories", c_uint64),
        ("kernel", c_uint64),
        ("kernel_ms", c_double),
        ("tail", c_uint64),
        ("capped", c_uint64),
    ]


class TuneResult(ctypes.Structure):
    """``rebop_tune_result`` of include/rebop_gpu.h."""

    _fields_ = [
        ("kernel", ctypes.c_int32),
        ("threads", ctypes.c_int32),
        ("occupancy", c_double),
        ("ms", c_double),
        ("n_tried", c_uint32),
    ]


class Job(ctypes.Structure):
    """``rebop_job`` of include/rebop_gpu.h."""

    _fields_ = [
        ("net", c_void_p),
        ("tmax", c_double),
        ("nb_steps", c_uint64),
        ("traj_begin", c_uint64),
        ("n_traj", c_uint64),
        ("d_x0", c_void_p),
        ("d_out", c_void_p),
        ("d_final", c_void_p),
        ("cost_hint", c_double),
    ]


class RebopError(RuntimeError):
    """A failure reported by the device library."""

    def __init__(self, code: int, message: str):
        super().__init__(message)
        self.code = code


class RebopValueError(RebopError, ValueError):
    """Invalid shapes, indices or values (the reference panics or raises ValueError)."""


class RebopCapError(RebopError):
    """Some trajectory reached the draw cap (the watchdog); their results stop there."""


_lib = None


def load() -> ctypes.CDLL:
    """Load the library once; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    path = Path(os.environ.get("REBOP_AMD_LIB", LIB_PATH))
    # PyTorch-ROCm ships its own libamdhip64 and loads it by file name; the
    # library binds libamdhip64.so.7 by SONAME.  Loading torch first makes
    # both use torch's copy: one HIP runtime per process.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    if not path.exists():
        raise RebopError(
            REBOP_ERR_DEVICE,
            f"{path} is missing: build it with `python -m rebop_amd.build` "
            "(there is no CPU fallback)",
        )
    lib = ctypes.CDLL(str(path))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(code: int) -> None:
    """Raise the matching exception for a non-zero status."""
    if code == REBOP_OK:
        return
    msg = load().rebop_last_error().decode(errors="replace")
    if code in (REBOP_ERR_SHAPE, REBOP_ERR_INDEX, REBOP_ERR_VALUE, REBOP_ERR_RANGE):
        raise RebopValueError(code, msg)
    if code == REBOP_ERR_CAP:
        raise RebopCapError(code, msg)
    raise RebopError(code, msg)


def device_info() -> tuple[str, int]:
    lib = load()
    buf = ctypes.create_string_buffer(128)
    cus = c_int(0)
    check(lib.rebop_device_info(buf, 128, ctypes.byref(cus)))
    return buf.value.decode(), cus.value
