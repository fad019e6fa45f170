"""ctypes binding of libgpucpm.so (the C ABI declared in include/gpucpm.h).

The library is required: importing this module raises ImportError when the
HIP build is missing, so no CPU fallback can stand in for it.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(_HERE, "libgpucpm.so")

CPM_OK = 0
CPM_EINVAL = -1
CPM_EHIP = -2
CPM_ESTATE = -3
CPM_ENOTIMPL = -4
NTYPES = 20


class RunArgs(ctypes.Structure):
    """Mirror of cpm_run_args (keyword arguments of Cpm.run)."""

    _fields_ = [(name, ctypes.c_int) for name in (
        "cell_sync", "block_sync", "global_sync", "threads_per_block",
        "positions_per_thread", "positions_per_checkerboard",
        "updates_per_checkerboard_switch", "updates_per_barrier", "iterations",
        "inner_iterations", "shared", "partial_dispatch", "centroid_tracking")]


# (name, restype, argtypes) of every entry point in include/gpucpm.h
_H = ctypes.c_void_p
_I = ctypes.c_int
_P = ctypes.POINTER
SIGNATURES = [
    ("cpm_last_error", ctypes.c_char_p, []),
    ("cpm_version", ctypes.c_char_p, []),
    ("cpm_create", _I, [_I, _I, _I, _I, _I, _I, _I, _P(_H)]),
    ("cpm_destroy", _I, [_H]),
    ("cpm_set_temperature", _I, [_H, ctypes.c_double]),
    ("cpm_set_area", _I, [_H, _I, _I, ctypes.c_double]),
    ("cpm_set_perimeter", _I, [_H, _I, _I, ctypes.c_double]),
    ("cpm_set_adhesion", _I, [_H, _I, _I, _I, _I]),
    ("cpm_set_stickiness", _I, [_H, _I, _I, _I]),
    ("cpm_set_fixed", _I, [_H, _I, _I]),
    ("cpm_add_cell", _I, [_H, _I, _I, _I, _I]),
    ("cpm_host_lattice", _I, [_H, _P(_P(ctypes.c_uint32))]),
    ("cpm_active_cells", _I, [_H, _P(_I)]),
    ("cpm_push", _I, [_H]),
    ("cpm_run", _I, [_H, _P(RunArgs)]),
    ("cpm_sync", _I, [_H]),
    ("cpm_pull", _I, [_H]),
    ("cpm_get_centroids", _I, [_H, _P(ctypes.c_double), _I, _P(_I)]),
    ("cpm_get_cell_table", _I, [_H, _P(ctypes.c_int32), _P(ctypes.c_int32), _P(ctypes.c_int32),
                                _P(ctypes.c_int64)]),
    ("cpm_set_autotune", _I, [_H, _I]),
    ("cpm_set_counters", _I, [_H, _I]),
    ("cpm_get_counters", _I, [_H, ctypes.POINTER(ctypes.c_int64), _I]),
    ("cpm_batch_set_autotune", _I, [_H, _I]),
    ("cpm_set_rng", _I, [_H, _I]),
    ("cpm_get_rng", _I, [_H, _P(_I)]),
    ("cpm_set_seed", _I, [_H, ctypes.c_uint64]),
    ("cpm_get_seed", _I, [_H, _P(ctypes.c_uint64)]),
    ("cpm_set_mcs", _I, [_H, ctypes.c_uint64]),
    ("cpm_get_mcs", _I, [_H, _P(ctypes.c_uint64)]),
    ("cpm_timing_begin", _I, [_H]),
    ("cpm_timing_end", _I, [_H, _P(ctypes.c_double), _P(ctypes.c_int64), _P(ctypes.c_int64)]),
    ("cpm_last_launch", _I, [_H, _P(_I), _P(_I), _P(_I)]),
    ("cpm_last_kernel", ctypes.c_char_p, [_H]),
    ("cpm_set_chemotaxis", _I, [_H, _I, ctypes.c_double]),
    ("cpm_host_field", _I, [_H, _P(_P(ctypes.c_double))]),
    ("cpm_set_act", _I, [_H, _I, _I, ctypes.c_double]),
    ("cpm_host_act", _I, [_H, _P(_P(ctypes.c_uint32))]),
    ("cpm_set_persistence", _I, [_H, _I, _I, ctypes.c_double, ctypes.c_double]),
    ("cpm_get_pref_dirs", _I, [_H, _P(ctypes.c_double)]),
    ("cpm_set_pref_dirs", _I, [_H, _P(ctypes.c_double)]),
    ("cpm_batch_create", _I, [_P(_H), _I, _P(_H)]),
    ("cpm_batch_run", _I, [_H, _P(RunArgs)]),
    ("cpm_batch_sync", _I, [_H]),
    ("cpm_batch_info", _I, [_H, _P(_I), _P(_I)]),
    ("cpm_batch_destroy", _I, [_H]),
    ("cpm_set_adhesion_real", _I, [_H, _I, _I, ctypes.c_double, _I]),
    ("cpm_energy_stats", _I, [_H, _P(ctypes.c_int64), _P(ctypes.c_int64)]),
    ("cpm_energy_stats_many", _I, [_P(_H), _I, _P(ctypes.c_int64), _P(ctypes.c_int64)]),
    ("cpm_energy_grad", _I, [_H, _I, _P(ctypes.c_double), _P(ctypes.c_double), _P(ctypes.c_double)]),
    ("cpm_set_launch", _I, [_H, _I, _I]),
    ("cpm_batch_set_blocks", _I, [_H, ctypes.c_double]),
    ("cpm_batch_set_form", _I, [_H, _I]),
    ("cpm_set_kernel_options", _I, [_H, _I, _I, _I, _I, _I, _I]),
    ("cpm_batch_set_kernel_options", _I, [_H, _I, _I, _I]),
    ("cpm_energy_stats_timing", _I, [_P(ctypes.c_double)]),
    ("cpm_batch_timing_begin", _I, [_H]),
    ("cpm_batch_timing_end", _I, [_H, _P(ctypes.c_double), _P(ctypes.c_int64), _P(ctypes.c_int64)]),
    ("cpm_batch_last_kernel", ctypes.c_char_p, [_H]),
    ("cpm_batch_graph_info", _I, [_H, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int64),
                                  ctypes.POINTER(ctypes.c_int64)]),
    ("cpm_energy_stats_many_noalloc", _I, [_P(_H), _I, _P(ctypes.c_int64), _P(ctypes.c_int64)]),
    ("cpm_resident_fallbacks", _I, [_H, _P(ctypes.c_int64), _P(ctypes.c_int64)]),
]

if not os.path.exists(LIBPATH):
    raise ImportError(
        f"gpucpm: {LIBPATH} is missing; build it with `python gpu-cpm_amd/build.py` "
        "(HIP for gfx950).  There is no CPU fallback.")

lib = ctypes.CDLL(LIBPATH)
for _name, _res, _args in SIGNATURES:
    _f = getattr(lib, _name)
    _f.restype = _res
    _f.argtypes = _args


def check(rc: int) -> None:
    """Raise the Python exception matching a negative CPM_E* code."""
    if rc == CPM_OK:
        return
    msg = (lib.cpm_last_error() or b"").decode()
    if rc == CPM_EINVAL:
        raise ValueError(msg)
    if rc == CPM_ENOTIMPL:
        raise NotImplementedError(msg)
    raise RuntimeError(msg)
