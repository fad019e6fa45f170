"""C ABI and Python surface checks that need no GPU.

- libgpucpm.so loads and exports every entry point include/gpucpm.h declares;
- argument validation happens before any device call (so it is testable
  here), and without a device cpm_create fails with CPM_EHIP, not a crash;
- gpucpm.Cpm keeps the reference's names and positional/keyword rules
  (python_wrapper.cpp:27-41, 55-129, 131-148, 182-227, 365-380).
"""
from __future__ import annotations

import ctypes
import inspect
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "gpucpm.h")


def _declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\**\s+\**(cpm_\w+)\s*\(", text, re.M)))


def test_header_declares_the_abi():
    names = _declared()
    for n in ("cpm_create", "cpm_destroy", "cpm_push", "cpm_run", "cpm_pull", "cpm_sync",
              "cpm_get_centroids", "cpm_set_adhesion", "cpm_set_area", "cpm_set_perimeter",
              "cpm_add_cell", "cpm_host_lattice", "cpm_last_error"):
        assert n in names


def test_library_exports_every_declared_symbol():
    import gpucpm
    lib = ctypes.CDLL(gpucpm._lib.LIBPATH)
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing
    bound = {s[0] for s in gpucpm._lib.SIGNATURES}
    assert set(_declared()) == bound


def _create(*args):
    from gpucpm._lib import lib
    h = ctypes.c_void_p()
    rc = lib.cpm_create(*args, ctypes.byref(h))
    return rc, lib.cpm_last_error().decode()


@pytest.mark.parametrize("args,needle", [
    ((100, 2, 10, 20, 0, 1, 1), "power of two"),
    ((64, 4, 10, 20, 0, 1, 1), "dimensionality"),
    ((64, 2, -1, 20, 0, 1, 1), "number_of_cells"),
    ((64, 2, 1 << 24, 20, 0, 1, 1), "number_of_cells"),
    ((2048, 3, 10, 20, 0, 1, 1), "too large"),
    ((64, 2, 10, 20, 0, 0, 1), "history"),
])
def test_create_validates_before_touching_the_device(args, needle):
    rc, msg = _create(*args)
    assert rc == -1 and needle in msg


def test_create_without_device_reports_hip_error():
    import torch  # noqa: F401  (only to learn whether a device is present)
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    rc, msg = _create(64, 2, 10, 20, 0, 1, 1)
    assert rc == -2 and "device" in msg


def test_null_handle_is_rejected():
    from gpucpm._lib import lib
    assert lib.cpm_push(None) == -1
    assert lib.cpm_sync(None) == -1


def test_python_constructor_is_positional_only():
    import gpucpm
    sig = inspect.signature(gpucpm.Cpm.__init__)
    params = list(sig.parameters.values())[1:]
    assert [p.name for p in params] == ["dimension", "dimensionality", "number_of_cells",
                                        "temperature", "has_field", "history", "has_act"]
    assert all(p.kind == p.POSITIONAL_ONLY for p in params)
    assert params[5].default == 1 and params[6].default is True
    with pytest.raises(TypeError):
        gpucpm.Cpm(dimension=64, dimensionality=2, number_of_cells=1, temperature=1, has_field=False)
    with pytest.raises(TypeError):
        gpucpm.Cpm(64.0, 2, 1, 1, False)  # 'i' refuses floats


def test_set_constraints_keywords_match_reference():
    import gpucpm
    sig = inspect.signature(gpucpm.Cpm.set_constraints)
    params = list(sig.parameters.values())[1:]
    assert params[0].name == "cell_type" and params[0].kind == params[0].POSITIONAL_OR_KEYWORD
    kw = [p.name for p in params[1:]]
    assert kw == ["other_cell_type", "lambda_perimeter", "target_perimeter", "lambda_area",
                  "target_area", "lambda_act", "max_act", "adhesion", "fixed",
                  "lambda_chemotaxis", "lambda_persistence", "persistence_diffusion",
                  "persistence_time", "set_adhesion_symmetric", "stickiness"]
    assert all(p.kind == p.KEYWORD_ONLY for p in params[1:])
    assert all(p.default == -1 for p in params[1:] if p.name != "set_adhesion_symmetric")


def test_run_signature_matches_reference():
    import gpucpm
    sig = inspect.signature(gpucpm.Cpm.run)
    params = list(sig.parameters.values())[1:]
    names = [p.name for p in params]
    assert names[:13] == ["cell_sync", "block_sync", "global_sync", "threads_per_block",
                          "positions_per_thread", "positions_per_checkerboard",
                          "updates_per_checkerboard_switch", "updates_per_barrier",
                          "iterations", "inner_iterations", "shared", "partial_dispatch",
                          "centroid_tracking"]
    assert all(p.kind == p.POSITIONAL_OR_KEYWORD for p in params[:13])
    assert [p.default for p in params[10:13]] == [0, 0, False]
    assert params[13].name == "seed" and params[13].kind == params[13].KEYWORD_ONLY


def test_method_table_matches_reference():
    import gpucpm
    for m in ("run", "push_to_gpu", "pull_from_gpu", "get_state", "get_field", "add_cell",
              "get_centroids", "synchronize", "get_act_state", "set_constraints"):
        assert callable(getattr(gpucpm.Cpm, m))
    sig = inspect.signature(gpucpm.Cpm.add_cell)
    params = list(sig.parameters.values())[1:]
    assert [p.name for p in params] == ["type", "x", "y", "z"]
    assert all(p.kind == p.POSITIONAL_ONLY for p in params) and params[3].default == -1
