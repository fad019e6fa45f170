"""Replays of the reference's example scripts (examples/example2d.py,
example3d.py, sorting.py) through gpucpm, minus the imaging, as
API-conformance tests; each is also checked bit-exact against the oracle
over a prefix of its run (same seed)."""
from __future__ import annotations

import numpy as np
import pytest

from helpers import assert_tables_equal

pytestmark = pytest.mark.gpu


def _oracle_from_sim(sim, L, d, ncells, T, J, lam_area, target_area, c, u, seed, nmcs, lattice,
                     ppt):
    """The scripts pass no ``shared``: the reference's default, non-shared
    roles with positions_per_thread ppt."""
    import oracle as O
    p = O.Params.make(T, J=J, lam_area=lam_area, target_area=target_area)
    lat = lattice.astype(np.uint32).copy()
    tab = O.tally(lat, d, ncells)
    O.sweep(lat, d, tab, p, c, u, seed, 0, nmcs, 8, shared=False, ppt=ppt)
    return lat, tab


def test_example2d_replay():
    """examples/example2d.py: 3 cells, area 200, adhesion 10, 2000 MCS."""
    import gpucpm
    print("CASE example2d", flush=True)
    dimension = 256
    sim = gpucpm.Cpm(dimension, 2, 10, 20, False)
    sim.add_cell(1, 100, 100)
    sim.add_cell(1, 200, 200)
    sim.add_cell(1, 200, 100)
    sim.set_constraints(cell_type=1, lambda_area=10, target_area=200)
    sim.set_constraints(cell_type=0, other_cell_type=1, adhesion=10)
    initial = sim.get_state().copy()
    sim.push_to_gpu()
    sim.set_seed(7)
    kw = dict(cell_sync=0, block_sync=0, global_sync=1, threads_per_block=4,
              positions_per_thread=8, positions_per_checkerboard=4,
              updates_per_checkerboard_switch=1, updates_per_barrier=1)
    sim.run(**kw, iterations=100, inner_iterations=100)
    sim.synchronize()
    sim.pull_from_gpu()
    state = sim.get_state().view(np.uint32).copy()
    J = np.zeros((2, 2), int)
    J[0, 1] = J[1, 0] = 10
    lat_o, tab_o = _oracle_from_sim(sim, dimension, 2, 10, 20, J, [0, 10], [0, 200], 4, 1, 7, 100,
                                    initial, ppt=8)
    np.testing.assert_array_equal(state, lat_o)
    assert_tables_equal(sim.get_cell_table(), tab_o, 2)
    # continue to the example's 2000 MCS and check what the example prints
    sim.run(**kw, iterations=1900, inner_iterations=1900)
    sim.synchronize()
    centroids = sim.get_centroids()
    assert centroids.shape == (3, 2)
    area = sim.get_cell_table()["area"][1:4]
    assert np.all(area > 120) and np.all(area < 260)  # grew from 1 site towards 200
    for (x, y), c in zip(centroids, [(100, 100), (200, 200), (200, 100)]):
        assert abs(x - c[0]) < 15 and abs(y - c[1]) < 15
    sim.pull_from_gpu()
    state = sim.get_state()
    cell_ids = state % 2**24
    assert set(np.unique(cell_ids)) == {0, 1, 2, 3}
    assert np.count_nonzero(cell_ids) == area.sum()


def test_example3d_replay():
    """examples/example3d.py: 64^3, 3 cells, volume 2000 (prefix of 60 MCS)."""
    import gpucpm
    print("CASE example3d", flush=True)
    dimension = 64
    sim = gpucpm.Cpm(dimension, 3, 10, 20, False)
    for x, y, z in [(16, 16, 16), (48, 48, 32), (32, 16, 48)]:  # the script's three seeds
        sim.add_cell(1, x, y, z)
    sim.set_constraints(cell_type=1, lambda_area=10, target_area=2000)
    sim.set_constraints(cell_type=0, other_cell_type=1, adhesion=10)
    initial = sim.get_state().copy()
    sim.push_to_gpu()
    sim.set_seed(11)
    sim.run(cell_sync=0, block_sync=0, global_sync=1, threads_per_block=4,
            positions_per_thread=8, positions_per_checkerboard=4,
            updates_per_checkerboard_switch=1, updates_per_barrier=1,
            iterations=60, inner_iterations=60)
    sim.synchronize()
    sim.pull_from_gpu()
    state = sim.get_state().view(np.uint32).copy()
    J = np.zeros((2, 2), int)
    J[0, 1] = J[1, 0] = 10
    lat_o, tab_o = _oracle_from_sim(sim, dimension, 3, 10, 20, J, [0, 10], [0, 2000], 4, 1, 11, 60,
                                    initial, ppt=8)
    np.testing.assert_array_equal(state, lat_o)
    assert_tables_equal(sim.get_cell_table(), tab_o, 3)
    assert sim.get_centroids().shape == (3, 3)
    assert np.all(sim.get_cell_table()["area"][1:4] > 100)


def test_sorting_example_replay():
    """examples/sorting.py: state written through get_state(), 1000 MCS with
    its schedule (c = 2, u = 1); a 200-MCS prefix is checked against the oracle.
    The lattice is this suite's own description of the script's set-up
    (stats_common.sorting_lattice: 3x3 cells on a 6-site grid inside a disc
    of radius L/3, two types at random)."""
    import gpucpm
    from stats_common import sorting_lattice
    print("CASE sorting.py", flush=True)
    dimension = 256
    lattice, ncells, _, _ = sorting_lattice(L=dimension, seed=5)
    sim = gpucpm.Cpm(dimension, 2, ncells, 10, False)
    sim.set_constraints(cell_type=1, lambda_area=1, target_area=40)
    sim.set_constraints(cell_type=1, other_cell_type=1, adhesion=14)
    sim.set_constraints(cell_type=0, other_cell_type=1, adhesion=16)
    sim.set_constraints(cell_type=2, lambda_area=1, target_area=40)
    sim.set_constraints(cell_type=2, other_cell_type=2, adhesion=2)
    sim.set_constraints(cell_type=2, other_cell_type=1, adhesion=11)
    sim.set_constraints(cell_type=2, other_cell_type=0, adhesion=16)
    sim.get_state()[...] = lattice.view(np.int32)
    sim.push_to_gpu()
    sim.set_seed(3)
    kw = dict(cell_sync=0, block_sync=0, global_sync=1, threads_per_block=32,
              positions_per_thread=4, positions_per_checkerboard=2,
              updates_per_checkerboard_switch=1, updates_per_barrier=1)
    sim.run(**kw, iterations=200, inner_iterations=200)
    sim.synchronize()
    sim.pull_from_gpu()
    got = sim.get_state().view(np.uint32).copy()
    J = np.array([[0, 16, 16], [16, 14, 11], [16, 11, 2]])
    lat_o, tab_o = _oracle_from_sim(sim, dimension, 2, ncells, 10, J, [0, 1, 1], [0, 40, 40], 2, 1,
                                    3, 200, lattice, ppt=4)
    np.testing.assert_array_equal(got, lat_o)
    assert_tables_equal(sim.get_cell_table(), tab_o, 2)
    sim.run(**kw, iterations=800, inner_iterations=800)
    sim.synchronize()
    sim.pull_from_gpu()
    sim_state = sim.get_state()
    types = sim_state // 2**24
    ids = sim_state % 2**24
    assert set(np.unique(types)) <= {0, 1, 2}
    area = sim.get_cell_table()["area"][1:]
    assert abs(area[area > 0].mean() - 40) < 12  # cells relax from 9 sites towards 40
    assert ids.max() <= ncells
