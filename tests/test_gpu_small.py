"""The single-workgroup form (launch form 3: lattice, cell tables and every
colour phase of a chunk in one workgroup's LDS) against the oracle,
bit-exact: both roles, several (c, u) schedules, fixed walls, chemotaxis,
3D persistence, 32-bit sites, more than four types, and the notebook's
interactive pattern (one MCS and a readback per frame,
examples/gpucpm_sorting_example.ipynb:316), which chooses this form by
itself."""
from __future__ import annotations

import numpy as np
import pytest

from helpers import assert_tables_equal, make_cpm, run_kwargs, run_oracle

pytestmark = pytest.mark.gpu


def _announce(*a):
    print("CASE", *a, flush=True)


def _six_types(L, seed):
    from gpucpm import workloads as W
    rng = np.random.default_rng(seed)
    lat = np.zeros((L, L), np.uint32)
    n = 0
    for y in range(0, L, 6):
        for x in range(0, L, 6):
            n += 1
            lat[y:y + 4, x:x + 4] = (np.uint32(rng.integers(1, 6)) << np.uint32(24)) | np.uint32(n)
    cons = [dict(cell_type=t, lambda_area=2, target_area=16) for t in range(1, 6)]
    cons += [dict(cell_type=t, other_cell_type=0, adhesion=8 + t) for t in range(1, 6)]
    cons += [dict(cell_type=1, other_cell_type=t, adhesion=3 * t) for t in range(1, 6)]
    return W.Workload("six_types", L, 2, n, 10, lat, cons)


def _cases():
    from gpucpm import workloads as W
    return [  # name, workload, c, u, shared, ppt, nmcs, inner, narrow
        ("2d-sh-c1", W.sorting2d(L=128, grid=9, edge=10, seed=201), 1, 1, True, 2, 20, 20, 1),
        ("2d-sh-c2u3", W.sorting2d(L=128, grid=9, edge=10, seed=202), 2, 3, True, 4, 12, 12, 1),
        ("2d-ns-notebook", W.sorting2d(L=256, grid=20, edge=10, seed=203), 2, 1, False, 4, 6, 6, 1),
        ("2d-ns-c1", W.sorting2d(L=128, grid=9, edge=10, seed=204), 1, 1, False, 2, 10, 10, 1),
        ("2d-wall", W.walled(W.sorting2d(L=64, grid=5, edge=10, seed=205)), 2, 1, False, 4, 15, 15, 1),
        ("2d-chem", W.chemotactic(W.sorting2d(L=64, grid=5, edge=10, seed=206)), 1, 1, True, 2, 15, 15, 1),
        ("2d-wide", W.sorting2d(L=128, grid=9, edge=10, seed=207), 2, 1, True, 4, 12, 12, 0),
        ("2d-six", _six_types(64, 208), 1, 1, True, 2, 15, 15, 1),
        ("3d-sh", W.plain3d(L=32, ncells=60, seed=209), 1, 1, True, 2, 4, 4, 1),
        ("3d-ns", W.plain3d(L=32, ncells=60, seed=210), 2, 3, False, 4, 3, 3, 1),
        ("3d-pers", W.persistent(W.plain3d(L=32, ncells=40, seed=211), lam=(0.0, 200.0),
                                 diffusion=(0.0, 0.6)), 1, 1, True, 2, 6, 2, 1),
    ]


@pytest.mark.parametrize("idx", range(11))
def test_small_form_matches_oracle(idx, engine_options):
    name, wl, c, u, shared, ppt, n, inner, narrow = _cases()[idx]
    engine_options(mode="small", narrow=narrow)
    _announce("small", name, wl.L, wl.d, c, u, shared, ppt)
    sim = make_cpm(wl)
    sim.push_to_gpu()
    sim.set_seed(300 + idx)
    sim.run(**run_kwargs(c, u, n, inner, shared=int(shared), positions_per_thread=ppt))
    sim.synchronize()
    sim.pull_from_gpu()
    launch = sim.last_launch()
    assert launch["form"] == 3 and "k_small_lds" in launch["kernel"]
    assert ("unsigned short" in launch["kernel"]) == bool(narrow)
    lat_o, tab_o = run_oracle(wl, c, u, 300 + idx, n, inner=inner, shared=shared, ppt=ppt)
    np.testing.assert_array_equal(sim.get_state().view(np.uint32), lat_o)
    assert_tables_equal(sim.get_cell_table(), tab_o, wl.d)
    if hasattr(tab_o, "pref"):
        np.testing.assert_array_equal(sim.get_pref_dirs().reshape(-1), np.asarray(tab_o.pref).reshape(-1))
    assert np.count_nonzero(sim.get_state().view(np.uint32) != wl.lattice) > 0


def test_interactive_frames_choose_small_form_and_match_oracle():
    """run(iterations=1) + pull_from_gpu per frame with the notebook's schedule
    (shared=0, ppt 4, c 2) at 256^2: the engine picks the single-workgroup
    form by itself; 12 frames equal 12 MCS of the oracle."""
    from gpucpm import workloads as W
    wl = W.sorting2d(L=256, grid=20, edge=10, seed=212)
    _announce("interactive", wl.L)
    sim = wl.make_cpm()
    sim.push_to_gpu()
    sim.set_seed(77)
    kw = run_kwargs(2, 1, 1, 1, shared=0, positions_per_thread=4, partial_dispatch=1)
    for _ in range(12):
        sim.run(**kw)
        sim.pull_from_gpu()
    assert sim.last_launch()["form"] == 3
    lat_o, tab_o = run_oracle(wl, 2, 1, 77, 12, shared=False, ppt=4)
    np.testing.assert_array_equal(sim.get_state().view(np.uint32), lat_o)
    assert_tables_equal(sim.get_cell_table(), tab_o, 2)


def test_auto_spreads_large_phases_over_resident_tiles():
    """256^2 with c = 1 has 16384 attempts per phase: one CU would be slower
    than many, so the automatic choice runs the phases on resident tiles
    (launch form 4); with c = 2 and u = 4, which that form does not take,
    it launches once per phase."""
    from gpucpm import workloads as W
    wl = W.plain2d()
    _announce("auto", wl.L)
    sim = wl.make_cpm()
    sim.push_to_gpu()
    sim.run(**run_kwargs(1, 1, 2))
    assert sim.last_launch()["form"] == 4
    sim.run(**run_kwargs(2, 4, 2))
    assert sim.last_launch()["form"] == 0
