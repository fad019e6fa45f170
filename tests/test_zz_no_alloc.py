"""The allocation-free statistics entry point (cpm_energy_stats_many_noalloc,
through the C ABI) against cpm_energy_stats_many on the same engines,
bit-exact, on two calls in a row (the second with fewer, smaller engines:
the shared scratch keeps the first call's contents), and its size limit."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

from helpers import run_kwargs

pytestmark = pytest.mark.gpu


def _stats(fn, sims):
    from gpucpm._lib import check
    n = len(sims)
    pairs = np.zeros((n, 20, 20), np.int64)
    mom = np.zeros((n, 20, 5), np.int64)
    hs = (ctypes.c_void_p * max(n, 1))(*[s._h.value for s in sims])
    P = ctypes.POINTER(ctypes.c_int64)
    check(fn(ctypes.cast(hs, ctypes.POINTER(ctypes.c_void_p)), n, pairs.ctypes.data_as(P),
             mom.ctypes.data_as(P)))
    return pairs, mom


def _six_types(L, seed):
    """Random 4x4 cells of types 1..5 (more than four types: the per-pair
    LDS path of the statistics kernel)."""
    from gpucpm import workloads as W
    rng = np.random.default_rng(seed)
    lat = np.zeros((L, L), np.uint32)
    n = 0
    for y in range(0, L, 6):
        for x in range(0, L, 6):
            n += 1
            lat[y:y + 4, x:x + 4] = (np.uint32(rng.integers(1, 6)) << np.uint32(24)) | np.uint32(n)
    cons = [dict(cell_type=t, lambda_area=2, target_area=16) for t in range(1, 6)]
    cons += [dict(cell_type=t, other_cell_type=0, adhesion=8) for t in range(1, 6)]
    return W.Workload("six_types", L, 2, n, 10, lat, cons)


def _engines(spec):
    from gpucpm import workloads as W
    sims = []
    for kind, L, seed, mcs in spec:
        if kind == "sort":
            wl = W.sorting2d(L=L, grid=(L - 8) // 10, edge=10, seed=seed)
        elif kind == "six":
            wl = _six_types(L, seed)
        else:
            wl = W.plain3d(L=L, ncells=40, seed=seed)
        s = wl.make_cpm()
        s.push_to_gpu()
        if mcs:  # a run leaves the 16-bit lattice current
            s.run(**run_kwargs(1, 1, mcs), seed=seed)
        sims.append(s)
    return sims


def test_noalloc_stats_equal_allocating_entry_point():
    from gpucpm._lib import lib
    first = [("sort", 128 if i % 3 else 64, 10 + i, 5 * (i % 2)) for i in range(34)]
    first += [("six", 64, 60 + i, 4 * i) for i in range(3)] + [("3d", 32, 70 + i, 2 * i) for i in range(3)]
    second = [("sort", 64, 80, 3), ("six", 32, 81, 0), ("3d", 32, 82, 1)]
    for name, spec in (("first", first), ("second", second)):
        print("CASE noalloc", name, len(spec), flush=True)
        sims = _engines(spec)
        p0, m0 = _stats(lib.cpm_energy_stats_many, sims)
        p1, m1 = _stats(lib.cpm_energy_stats_many_noalloc, sims)
        np.testing.assert_array_equal(p1, p0)
        np.testing.assert_array_equal(m1, m0)
        assert p0.sum() > 0 and m0[:, 1:, 0].sum() > 0


def test_noalloc_refuses_more_engines_than_its_scratch():
    from gpucpm._lib import lib
    from gpucpm import workloads as W
    print("CASE noalloc limit", flush=True)
    sims = []
    for i in range(65):
        s = W.sorting2d(L=16, grid=1, edge=10, seed=i).make_cpm()
        s.push_to_gpu()
        sims.append(s)
    with pytest.raises(ValueError, match="at most 64"):
        _stats(lib.cpm_energy_stats_many_noalloc, sims)
    p0, m0 = _stats(lib.cpm_energy_stats_many, sims[:64])
    p1, m1 = _stats(lib.cpm_energy_stats_many_noalloc, sims[:64])
    np.testing.assert_array_equal(p1, p0)
    np.testing.assert_array_equal(m1, m0)
