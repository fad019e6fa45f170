"""GPU parity of the 16-bit lattice that the per-phase kernels run on when
every site fits (cell | type << idbits, idbits = 16 - the bits of the
largest type), and of the 32-bit kernels otherwise.

Every case prints its name, shape and schedule before touching the GPU.
"""
from __future__ import annotations

import dataclasses

import numpy as np
import pytest

from helpers import assert_tables_equal, run_kwargs, run_oracle, make_cpm

pytestmark = pytest.mark.gpu


def _announce(*a):
    print("CASE", *a, flush=True)


def _tiny_cells(ncells):
    """256^2 sorting lattice of 2x2 cells numbered 1..ncells (types 1/2,
    three types in all: 14 cell bits in 16-bit sites)."""
    from gpucpm import workloads as W
    return W.sorting2d(L=256, grid=128, edge=2, seed=91, ncells=ncells)


def _retyped(wl, old, new):
    """The workload with cell type ``old`` renamed ``new`` (lattice and
    constraints)."""
    lat = wl.lattice.copy()
    m = (lat >> 24) == old
    lat[m] = (lat[m] & 0xFFFFFF) | np.uint32(new << 24)
    cons = []
    for kw in wl.constraints:
        kw = dict(kw)
        for key in ("cell_type", "other_cell_type"):
            if kw.get(key) == old:
                kw[key] = new
        cons.append(kw)
    return dataclasses.replace(wl, lattice=lat, constraints=cons)


def _run(wl, c, u, seed, splits, between=None):
    sim = make_cpm(wl)
    sim.push_to_gpu()
    sim.set_seed(seed)
    for i, n in enumerate(splits):
        if i and between:
            between(sim)
        sim.run(**run_kwargs(c, u, n))
    sim.synchronize()
    sim.pull_from_gpu()
    return sim, sim.get_state().view(np.uint32).copy(), sim.get_cell_table()


@pytest.mark.parametrize("ncells,narrow", [(16383, True), (16384, False)])
def test_cell_bits_limit(ncells, narrow):
    """16383 cells of three types fit in 16-bit sites (the largest cell id
    uses all 14 cell bits); 16384 do not, and the run uses 32-bit sites.
    Both reproduce the oracle."""
    wl = _tiny_cells(ncells)
    _announce("limit", ncells, wl.L)
    sim, lat, table = _run(wl, 1, 1, 5, [6])
    kernel = sim.last_launch()["kernel"]
    assert ("unsigned short" in kernel) == narrow, kernel
    lat_o, tab_o = run_oracle(wl, 1, 1, 5, 6)
    np.testing.assert_array_equal(lat, lat_o)
    assert_tables_equal(table, tab_o, 2)


@pytest.mark.parametrize("newtype", [5, 15])
def test_large_types(newtype):
    """A type of 5 (3 type bits, 13 cell bits) or 15 (4 and 12) keeps results
    exact."""
    from gpucpm import workloads as W
    wl = _retyped(W.walled(W.sorting2d(L=128, grid=9, edge=10, seed=92)), 2, newtype)
    _announce("types", newtype, wl.L)
    sim, lat, table = _run(wl, 1, 1, 8, [20])
    assert "unsigned short" in sim.last_launch()["kernel"]
    lat_o, tab_o = run_oracle(wl, 1, 1, 8, 20)
    np.testing.assert_array_equal(lat, lat_o)
    assert_tables_equal(table, tab_o, 2)


def test_readback_between_runs():
    """Pulls and table reads between runs (which rebuild the 32-bit
    lattice) leave the run sequence unchanged."""
    from gpucpm import workloads as W
    wl = W.plain3d(L=32, ncells=60, seed=93)
    _announce("readback", wl.L)

    def read(sim):
        sim.pull_from_gpu()
        sim.get_cell_table()

    _, lat1, t1 = _run(wl, 1, 1, 9, [6])
    _, lat2, t2 = _run(wl, 1, 1, 9, [2, 3, 1], between=read)
    np.testing.assert_array_equal(lat1, lat2)
    for k in ("area", "perimeter", "csum"):
        np.testing.assert_array_equal(t1[k], t2[k])


def test_push_after_run_replaces_lattice():
    """A push after a run replaces both lattice copies: the next run starts
    from the pushed lattice."""
    from gpucpm import workloads as W
    wl = W.sorting2d(L=128, grid=9, edge=10, seed=94)
    _announce("repush", wl.L)
    sim = make_cpm(wl)
    sim.push_to_gpu()
    sim.set_seed(3)
    sim.run(**run_kwargs(1, 1, 10))
    sim.synchronize()
    sim.pull_from_gpu()
    edited = sim.get_state().view(np.uint32).copy()
    edited[:8, :8] = 0
    sim.get_state()[...] = edited.view(np.int32)
    sim.push_to_gpu()
    mcs = sim.mcs
    sim.run(**run_kwargs(1, 1, 10))
    sim.synchronize()
    sim.pull_from_gpu()
    lat = sim.get_state().view(np.uint32).copy()
    lat_o, tab_o = run_oracle(wl, 1, 1, 3, 10, mcs0=mcs, lattice=edited)
    np.testing.assert_array_equal(lat, lat_o)
    assert_tables_equal(sim.get_cell_table(), tab_o, 2)


def test_energy_stats_equal_on_both_lattices(engine_options):
    """The dH/dtheta statistics after a run on 16-bit sites equal those
    after the same run on 32-bit sites."""
    from gpucpm import workloads as W
    wl = W.sorting2d(L=128, grid=9, edge=10, seed=95)
    _announce("stats", wl.L)
    out = []
    for narrow in ("1", "0"):
        engine_options(narrow=narrow)
        sim, lat, _ = _run(wl, 1, 1, 4, [10])
        assert ("unsigned short" in sim.last_launch()["kernel"]) == (narrow == "1")
        out.append((lat, sim.energy_stats()))
    np.testing.assert_array_equal(out[0][0], out[1][0])
    for a, b in zip(out[0][1], out[1][1]):
        np.testing.assert_array_equal(a, b)
