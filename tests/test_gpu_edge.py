"""Edge cases of the sweep on the GPU against the oracle: the smallest
lattices, a lattice of medium only, one cell covering everything, the 16-bit
site limit with the largest type, and size-independent properties at the
largest sizes run here (4096^2 and 512^3: tables equal a recount).
Every case prints its name and shape before touching the GPU."""
from __future__ import annotations

import numpy as np
import pytest

from helpers import assert_tables_equal, run_kwargs, run_oracle

pytestmark = pytest.mark.gpu


def _announce(*a):
    print("CASE", *a, flush=True)


def _wl(lat, d, ncells, cons, T=10):
    from gpucpm.workloads import Workload
    return Workload("edge", lat.shape[0], d, ncells, T, np.ascontiguousarray(lat, np.uint32), cons)


def _cons(types):
    cons = []
    for t in types:
        cons.append(dict(cell_type=t, lambda_area=2, target_area=3))
        cons.append(dict(cell_type=t, other_cell_type=0, adhesion=4))
        for o in types:
            if o <= t:
                cons.append(dict(cell_type=t, other_cell_type=o, adhesion=2 + (t + o) % 5))
    return cons


def _random_lattice(L, d, ncells, types, seed):
    rng = np.random.default_rng(seed)
    ids = rng.integers(0, ncells + 1, size=(L,) * d).astype(np.uint32)
    ids[ids == 0] = 0
    t = np.asarray(types, np.uint32)[ids % len(types)]
    return np.where(ids > 0, (t << np.uint32(24)) | ids, 0).astype(np.uint32)


def _run(wl, c, u, seed, n, **extra):
    sim = wl.make_cpm()
    sim.push_to_gpu()
    sim.set_seed(seed)
    sim.run(**run_kwargs(c, u, n, **extra))
    sim.synchronize()
    sim.pull_from_gpu()
    return sim, sim.get_state().view(np.uint32).copy(), sim.get_cell_table()


@pytest.mark.parametrize("d,L,c,u,shared,ppt", [(2, 2, 1, 1, 1, 2), (3, 2, 1, 1, 1, 2),
                                                (2, 4, 2, 4, 1, 4), (2, 4, 2, 3, 0, 4),
                                                (3, 4, 1, 1, 0, 2), (2, 8, 1, 1, 0, 8)])
def test_smallest_lattices(d, L, c, u, shared, ppt):
    lat = _random_lattice(L, d, 3, (1, 2), seed=L * 10 + d)
    wl = _wl(lat, d, 3, _cons((1, 2)))
    _announce("small", d, L, c, u, shared, ppt)
    _, got, table = _run(wl, c, u, 5, 12, shared=shared, positions_per_thread=ppt)
    lat_o, tab_o = run_oracle(wl, c, u, 5, 12, shared=bool(shared), ppt=ppt)
    np.testing.assert_array_equal(got, lat_o)
    assert_tables_equal(table, tab_o, d)


@pytest.mark.parametrize("fill", ["medium", "one_cell"])
def test_uniform_lattices_do_not_change(fill):
    """With every site equal no attempt is active: nothing changes."""
    for d, L in ((2, 64), (3, 16)):
        lat = np.zeros((L,) * d, np.uint32) if fill == "medium" else np.full((L,) * d, (1 << 24) | 1, np.uint32)
        wl = _wl(lat, d, 1, _cons((1,)))
        _announce("uniform", fill, d, L)
        _, got, table = _run(wl, 1, 1, 7, 5)
        np.testing.assert_array_equal(got, lat)
        assert table["area"][1] == (0 if fill == "medium" else L ** d)
        assert table["perimeter"][1] == 0


@pytest.mark.parametrize("ncells,narrow", [(2047, True), (2048, False)])
def test_largest_type_and_cell_limit(ncells, narrow):
    """Type 19 present (5 type bits): 2047 cells fit in 16-bit sites, 2048
    do not; both reproduce the oracle."""
    from gpucpm import workloads as W
    base = W.sorting2d(L=128, grid=46, edge=2, seed=93, ncells=ncells)
    lat = base.lattice.copy()
    m = (lat >> 24) == 2
    lat[m] = (lat[m] & 0xFFFFFF) | np.uint32(19 << 24)
    wl = _wl(lat, 2, ncells, _cons((1, 19)))
    _announce("type19", ncells)
    sim, got, table = _run(wl, 1, 1, 9, 10)
    assert ("unsigned short" in sim.last_launch()["kernel"]) == narrow
    lat_o, tab_o = run_oracle(wl, 1, 1, 9, 10)
    np.testing.assert_array_equal(got, lat_o)
    assert_tables_equal(table, tab_o, 2)


@pytest.mark.parametrize("d,L,ncells,n", [(2, 4096, 60000, 20), (3, 512, 60000, 3)])
def test_largest_sizes_tables_equal_recount(d, L, ncells, n):
    """Size-independent property at 4096^2 and 512^3: after the run the
    incrementally kept tables equal a recount of the final lattice, every
    site holds a valid value, and the lattice changed."""
    import oracle as O
    from gpucpm import workloads as W
    wl = W.plain2d(L=L, ncells=ncells, seed=95) if d == 2 else W.plain3d(L=L, ncells=ncells, seed=96)
    _announce("largest", d, L, ncells, n)
    _, got, table = _run(wl, 1, 1, 11, n)
    rec = O.tally(got, d, wl.ncells)
    np.testing.assert_array_equal(table["area"], rec.area)
    np.testing.assert_array_equal(table["perimeter"], rec.perim)
    assert int(got.max() & 0xFFFFFF) <= wl.ncells
    assert np.count_nonzero(got != wl.lattice) > 0
