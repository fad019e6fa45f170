"""The resident form (launch form 4: 2D tiles resident in LDS for a launch,
holding the rows of their cells and exchanging border sites and shared
cells' deltas with their eight neighbours between colour phases) against
the oracle, bit-exact; the guard (a cell wider than two tiles, or met by a
tile not holding it) sends a launch back to one launch per phase, still
exact, and shortens the following launches."""
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


def _run(wl, seed, n, inner, narrow=1, tile=0):
    from helpers import ENGINE_OPTIONS
    ENGINE_OPTIONS.update(mode="resident", narrow=narrow, tile=tile)
    try:
        sim = make_cpm(wl)
    finally:
        ENGINE_OPTIONS.clear()
    sim.push_to_gpu()
    sim.set_seed(seed)
    sim.run(**run_kwargs(1, 1, n, inner))
    sim.synchronize()
    sim.pull_from_gpu()
    return sim


def _cases():
    from gpucpm import workloads as W
    return [  # name, workload, nmcs, inner, narrow, tile edge (0: the engine's, 64 from 256^2)
        ("sorting1024", W.sorting2d(seed=401), 40, 20, 1, 0),
        ("sorting512", W.sorting2d(L=512, grid=40, edge=10, seed=402), 30, 10, 1, 32),
        ("plain256", W.plain2d(seed=403), 60, 30, 1, 0),
        # (cells that grow or drift by a few sites per ten MCS: launches of
        # 10 MCS stay within the tiles that hold them)
        ("small-cells256", W.sorting_small_cells(L=256, seed=404), 40, 10, 1, 32),
        ("wall-chem256", W.chemotactic(W.walled(W.sorting2d(L=256, grid=20, edge=10, seed=405)),
                                       lam=(0.0, 20.0, 5.0, 0.0)), 30, 15, 1, 32),
        ("wide512", W.sorting2d(L=512, grid=40, edge=10, seed=406), 20, 20, 0, 0),
        ("six-types256", _six_types(256, 407), 30, 10, 1, 32),
        ("plain128", W.plain2d(L=128, ncells=60, seed=411), 30, 10, 1, 0),
    ]


@pytest.mark.parametrize("idx", range(8))
def test_resident_matches_oracle(idx):
    name, wl, n, inner, narrow, tile = _cases()[idx]
    _announce("resident", name, wl.L, n, inner, narrow, tile)
    sim = _run(wl, 500 + idx, n, inner, narrow, tile)
    launch = sim.last_launch()
    lat_o, tab_o = run_oracle(wl, 1, 1, 500 + idx, n)
    np.testing.assert_array_equal(sim.get_state().view(np.uint32), lat_o)
    assert_tables_equal(sim.get_cell_table(), tab_o, 2)
    guard, waits = sim.resident_fallbacks
    assert waits == 0
    # the frame is one cell around the whole lattice: the guard sends it
    # back; elsewhere the form must hold
    if name != "wall-chem256":
        assert guard == 0 and launch["form"] == 4, (guard, launch)
        assert "k_resident2d" in launch["kernel"]
        assert launch["tile"] == (tile or (64 if wl.L >= 256 else 32)), launch


def _general_cases():
    from gpucpm import workloads as W
    return [  # name, workload, c, shared, positions_per_thread, nmcs, inner, narrow, tile edge
        # the reference sorting.py schedule (non-shared roles, c 2, ppt 4)
        ("sortingpy1024", W.sorting2d(seed=421), 2, 0, 4, 8, 8, 1, 0),
        ("sortingpy512", W.sorting2d(L=512, grid=40, edge=10, seed=422), 2, 0, 4, 20, 20, 1, 32),
        ("ns-c1-256", W.sorting2d(L=256, grid=20, edge=10, seed=423), 1, 0, 2, 20, 10, 1, 32),
        ("ns-c4-256", W.plain2d(seed=424), 4, 0, 8, 20, 20, 1, 0),
        ("shared-c2-256", W.sorting2d(L=256, grid=20, edge=10, seed=425), 2, 1, 4, 20, 20, 1, 32),
        ("chem-ns-c2-256", W.chemotactic(W.sorting2d(L=256, grid=20, edge=10, seed=426), lam=(0.0, 20.0, 5.0)),
         2, 0, 4, 20, 20, 1, 0),
        ("six-types-ns-256", _six_types(256, 427), 2, 0, 4, 20, 10, 1, 32),
        ("wide-ns-512", W.sorting2d(L=512, grid=40, edge=10, seed=428), 2, 0, 4, 10, 10, 0, 0),
        ("shared-c2-256-t64", W.sorting2d(L=256, grid=20, edge=10, seed=429), 2, 1, 4, 20, 20, 1, 64),
        ("ns-c4-256-t32", W.plain2d(seed=430), 4, 0, 8, 20, 20, 1, 32),
    ]


@pytest.mark.parametrize("idx", range(10))
def test_resident_general_schedules_match_oracle(idx):
    """Any schedule with u = 1: each tile makes the attempts whose target it
    holds (with non-shared roles the target may be a sub-block's neighbour
    site in another tile), bit-exact against the oracle."""
    from helpers import ENGINE_OPTIONS
    name, wl, c, shared, ppt, n, inner, narrow, tile = _general_cases()[idx]
    _announce("resident general", name, wl.L, c, shared, ppt, n, inner, tile)
    ENGINE_OPTIONS.update(mode="resident", narrow=narrow, tile=tile)
    try:
        sim = make_cpm(wl)
    finally:
        ENGINE_OPTIONS.clear()
    sim.push_to_gpu()
    sim.set_seed(530 + idx)
    sim.run(**run_kwargs(c, 1, n, inner, shared=shared, positions_per_thread=ppt))
    sim.synchronize()
    sim.pull_from_gpu()
    launch = sim.last_launch()
    lat_o, tab_o = run_oracle(wl, c, 1, 530 + idx, n, shared=bool(shared), ppt=ppt)
    np.testing.assert_array_equal(sim.get_state().view(np.uint32), lat_o)
    assert_tables_equal(sim.get_cell_table(), tab_o, 2)
    guard, waits = sim.resident_fallbacks
    assert waits == 0
    assert guard == 0 and launch["form"] == 4, (guard, launch)
    assert launch["tile"] == (tile or 64), launch


@pytest.mark.parametrize("case", ["plain128", "sorting256"])
def test_many_short_launches_match_oracle(case):
    """Launches of one MCS each: every launch's first phase draws with the
    Philox blocks made at its set-up (with one attempt per thread, thread t
    reads the block of pair t / 2, which another wave wrote), bit-exact."""
    from gpucpm import workloads as W
    wl = W.plain2d(L=128, ncells=60, seed=411) if case == "plain128" else W.sorting2d(L=256, grid=20, edge=10,
                                                                                       seed=412)
    _announce("short launches", case, wl.L)
    n = 12
    sim = _run(wl, 515, n, 1)
    assert sim.last_launch()["form"] == 4
    lat_o, tab_o = run_oracle(wl, 1, 1, 515, n)
    np.testing.assert_array_equal(sim.get_state().view(np.uint32), lat_o)
    assert_tables_equal(sim.get_cell_table(), tab_o, 2)
    assert sim.resident_fallbacks == (0, 0)


def test_guard_sends_a_spread_cell_back_to_per_phase_launches():
    """One cell of 70 x 70 sites at 256^2 (tiles of 32): it spans three tiles,
    so the chunk is restored and run with one launch per phase, exactly."""
    from gpucpm import workloads as W
    wl = W.plain2d(seed=408)
    lat = wl.lattice.copy()
    lat[100:170, 60:130] = (1 << 24) | 1
    wl.lattice = lat
    _announce("guard", wl.L)
    sim = _run(wl, 509, 20, 10, tile=32)
    assert sim.resident_fallbacks[0] >= 1 and sim.resident_fallbacks[1] == 0
    lat_o, tab_o = run_oracle(wl, 1, 1, 509, 20)
    np.testing.assert_array_equal(sim.get_state().view(np.uint32), lat_o)
    assert_tables_equal(sim.get_cell_table(), tab_o, 2)


def test_dense_small_cells_move_to_smaller_tiles():
    """1849 cells of 16 sites at 256^2: a 64^2 tile with its margins meets
    more cells than its slots hold, so the first launch is redone per phase
    and the rest of the run (and the next run) takes 32^2 tiles, exactly."""
    wl = _six_types(256, 412)
    _announce("retile", wl.L)
    sim = _run(wl, 513, 30, 10)
    guard, waits = sim.resident_fallbacks
    assert guard == 1 and waits == 0, (guard, waits)
    assert sim.last_launch()["form"] == 4 and sim.last_launch()["tile"] == 32
    lat_o, tab_o = run_oracle(wl, 1, 1, 513, 30)
    np.testing.assert_array_equal(sim.get_state().view(np.uint32), lat_o)
    assert_tables_equal(sim.get_cell_table(), tab_o, 2)
    sim.run(**run_kwargs(1, 1, 10, 10))
    sim.synchronize()
    assert sim.resident_fallbacks[0] == 1 and sim.last_launch()["tile"] == 32


def test_launch_length_adapts_to_moving_cells():
    """Cells of the sorting example grow from 3 x 3 to ~40 sites and drift:
    one launch of 40 MCS meets cells outside the tiles holding them, so it is
    redone per phase and the next launches are shorter; every result stays
    bit-exact, and later runs of the engine keep to launches that hold."""
    from gpucpm import workloads as W
    wl = W.sorting_small_cells(L=256, seed=404)
    _announce("adaptive launches", wl.L)
    sim = _run(wl, 503, 40, 40, tile=32)
    guard, waits = sim.resident_fallbacks
    assert guard >= 1 and waits == 0
    lat_o, tab_o = run_oracle(wl, 1, 1, 503, 40)
    np.testing.assert_array_equal(sim.get_state().view(np.uint32), lat_o)
    assert_tables_equal(sim.get_cell_table(), tab_o, 2)
    # a second run of 40 MCS starts with the shortened launches
    sim.run(**run_kwargs(1, 1, 40, 40))
    sim.synchronize()
    sim.pull_from_gpu()
    assert sim.last_launch()["tile"] == 32
    lat_o2, tab_o2 = run_oracle(wl, 1, 1, 503, 80)
    np.testing.assert_array_equal(sim.get_state().view(np.uint32), lat_o2)
    assert_tables_equal(sim.get_cell_table(), tab_o2, 2)
    assert sim.last_launch()["form"] == 4


def test_resident_split_runs_equal_one_run():
    from gpucpm import workloads as W
    wl = W.sorting2d(L=512, grid=40, edge=10, seed=410)
    _announce("resident split", wl.L)
    a = _run(wl, 7, 24, 24).get_state().copy()
    from helpers import ENGINE_OPTIONS
    ENGINE_OPTIONS.update(mode="resident")
    try:
        sim = make_cpm(wl)
    finally:
        ENGINE_OPTIONS.clear()
    sim.push_to_gpu()
    sim.set_seed(7)
    for n, inner in ((8, 4), (16, 16)):
        sim.run(**run_kwargs(1, 1, n, inner))
    sim.synchronize()
    sim.pull_from_gpu()
    np.testing.assert_array_equal(sim.get_state(), a)
