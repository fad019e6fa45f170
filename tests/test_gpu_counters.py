"""The attempt counters (cpm_set_counters, the reference's analytics block
made live) against the oracle's counts of the same runs, exactly: attempts
drawn, attempts whose source and target differ, evaluated, accepted and
committed, for the compacted and uncompacted per-phase kernels, both roles,
cell_sync, 3D, fixed walls, the split act phases, the single-workgroup and
resident launch forms, and batched replicas (counted per engine).  The
counters change nothing: the lattice stays bit-exact with them on."""
from __future__ import annotations

import numpy as np
import pytest

from helpers import assert_tables_equal, run_kwargs, run_oracle

pytestmark = pytest.mark.gpu


def _cases():
    from gpucpm import workloads as W
    return [  # name, workload, c, u, shared, cell_sync, launch form (-1 auto), nmcs
        ("2d-compact", W.sorting2d(L=256, grid=20, edge=10, seed=601), 1, 1, 1, 0, 0, 20),
        ("2d-ns-c2u4", W.sorting2d(L=128, grid=9, edge=10, seed=602), 2, 4, 0, 0, 0, 10),
        ("2d-sync", W.sorting2d(L=128, grid=9, edge=10, seed=603), 1, 1, 1, 1, 0, 10),
        ("2d-sync-c2u4", W.sorting2d(L=128, grid=9, edge=10, seed=604), 2, 4, 0, 1, 0, 6),
        ("3d", W.plain3d(L=32, ncells=60, seed=605), 1, 1, 1, 0, 0, 6),
        ("2d-wall", W.walled(W.sorting2d(L=64, grid=5, edge=10, seed=606)), 1, 1, 1, 0, 0, 15),
        ("2d-act-split", W.active(W.sorting2d(L=64, grid=5, edge=10, seed=607)), 1, 1, 1, 0, 0, 10),
        ("2d-small", W.sorting2d(L=128, grid=9, edge=10, seed=608), 2, 1, 0, 0, 3, 8),
        ("2d-resident", W.sorting2d(L=256, grid=20, edge=10, seed=609), 1, 1, 1, 0, 4, 12),
        ("2d-resident-ns-c2", W.sorting2d(L=256, grid=20, edge=10, seed=610), 2, 1, 0, 0, 4, 10),
    ]


@pytest.mark.parametrize("idx", range(10))
def test_counters_equal_oracle(idx):
    name, wl, c, u, shared, sync, form, n = _cases()[idx]
    print("CASE counters", name, flush=True)
    sim = wl.make_cpm()
    if form >= 0:
        sim.set_launch(form, 0)
    sim.push_to_gpu()
    sim.set_seed(700 + idx)
    sim.set_counters(True)
    sim.run(**run_kwargs(c, u, n, shared=shared, cell_sync=sync, positions_per_thread=2 * c))
    sim.synchronize()
    got = sim.counters()
    if form >= 0:
        assert sim.last_launch()["form"] == form
    sim.pull_from_gpu()
    counts = np.zeros(4, np.int64)
    lat_o, tab_o = run_oracle(wl, c, u, 700 + idx, n, shared=bool(shared), cell_sync=bool(sync),
                              ppt=2 * c, counts=counts)
    np.testing.assert_array_equal(sim.get_state().view(np.uint32), lat_o)
    assert_tables_equal(sim.get_cell_table(), tab_o, wl.d)
    assert got["attempts"] == wl.L ** wl.d * n
    assert [got[k] for k in ("differing", "evaluated", "accepted", "committed")] == counts.tolist(), \
        (got, counts)
    assert got["committed"] > 0


def test_counters_per_engine_in_a_batch_and_reset():
    import gpucpm
    from gpucpm import workloads as W
    wls = W.batched_sorting(nrep=4, seed=61, sizes=(32, 64, 128))
    print("CASE counters batch", [w.L for w in wls], flush=True)
    sims = []
    for i, wl in enumerate(wls):
        s = wl.make_cpm()
        s.push_to_gpu()
        s.set_seed(800 + i)
        s.set_counters(True)
        sims.append(s)
    batch = gpucpm.Batch(sims)
    batch.run(**run_kwargs(1, 1, 12))
    batch.synchronize()
    for i, (wl, s) in enumerate(zip(wls, sims)):
        got = s.counters(reset=True)
        counts = np.zeros(4, np.int64)
        run_oracle(wl, 1, 1, 800 + i, 12, counts=counts)
        assert got["attempts"] == wl.L ** 2 * 12
        assert [got[k] for k in ("differing", "evaluated", "accepted", "committed")] == counts.tolist()
    # after a reset the counts start again; switched off they stop
    s = sims[0]
    s.run(**run_kwargs(1, 1, 3))
    s.synchronize()
    after = s.counters()
    assert after["attempts"] == wls[0].L ** 2 * 3 and after["differing"] > 0
    s.set_counters(False)
    s.run(**run_kwargs(1, 1, 3))
    s.synchronize()
    assert s.counters()["differing"] == after["differing"]


def test_counters_survive_redone_resident_launches():
    """A resident launch that fails its guard is redone per phase: the
    counters go back to their state before the launch and count the redone
    MCS once, so the totals still equal the oracle's."""
    from gpucpm import workloads as W
    wl = W.sorting_small_cells(L=256, seed=404)
    print("CASE counters redone launches", wl.L, flush=True)
    sim = wl.make_cpm()
    sim.set_launch(4, 32)
    sim.push_to_gpu()
    sim.set_seed(503)
    sim.set_counters(True)
    sim.run(**run_kwargs(1, 1, 40, 40))
    sim.synchronize()
    got = sim.counters()
    assert sim.resident_fallbacks[0] >= 1
    sim.pull_from_gpu()
    counts = np.zeros(4, np.int64)
    lat_o, tab_o = run_oracle(wl, 1, 1, 503, 40, counts=counts)
    np.testing.assert_array_equal(sim.get_state().view(np.uint32), lat_o)
    assert_tables_equal(sim.get_cell_table(), tab_o, 2)
    assert got["attempts"] == wl.L ** 2 * 40
    assert [got[k] for k in ("differing", "evaluated", "accepted", "committed")] == counts.tolist()
