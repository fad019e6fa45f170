"""Batched replicas on the GPU: one cooperative launch for many engines of
different sizes and parameters; each replica must end bit-identical to its
own single run and to the oracle."""
from __future__ import annotations

import numpy as np
import pytest

from helpers import assert_tables_equal, run_kwargs, run_oracle

pytestmark = pytest.mark.gpu


def _small_batch():
    from gpucpm import workloads as W
    # every other replica inside a fixed wall (per-replica fixed masks), one
    # with a chemotaxis field
    return [W.walled(w) if i % 2 else (W.chemotactic(w) if i == 2 else w)
            for i, w in enumerate(W.batched_sorting(nrep=6, seed=40, sizes=(32, 64, 128)))]


@pytest.mark.parametrize("form", [0, 1])
def test_batch_equals_single_runs_and_oracle(form):
    import gpucpm
    wls = _small_batch()
    print("CASE batch form", form, [w.L for w in wls], flush=True)
    sims = []
    for i, wl in enumerate(wls):
        s = wl.make_cpm()
        s.push_to_gpu()
        s.set_seed(100 + i)
        sims.append(s)
    batch = gpucpm.Batch(sims)
    batch.set_form(form)
    kw = run_kwargs(1, 1, 30, 10)
    batch.run(**kw)
    batch.synchronize()
    for i, (wl, s) in enumerate(zip(wls, sims)):
        s.pull_from_gpu()
        lat = s.get_state().view(np.uint32).copy()
        # the same engine alone
        t = wl.make_cpm()
        t.push_to_gpu()
        t.run(**kw, seed=100 + i)
        t.synchronize()
        t.pull_from_gpu()
        np.testing.assert_array_equal(lat, t.get_state().view(np.uint32))
        lat_o, tab_o = run_oracle(wl, 1, 1, 100 + i, 30)
        np.testing.assert_array_equal(lat, lat_o)
        assert_tables_equal(s.get_cell_table(), tab_o, 2)
        assert s.mcs == 30


def test_batch_continues_single_runs():
    """A batch run after single runs continues each replica's stream."""
    import gpucpm
    wls = _small_batch()[:3]
    print("CASE batch-continue", flush=True)
    sims = []
    for i, wl in enumerate(wls):
        s = wl.make_cpm()
        s.push_to_gpu()
        s.set_seed(7 + i)
        s.run(**run_kwargs(2, 4, 10))
        sims.append(s)
    gpucpm.Batch(sims).run(**run_kwargs(2, 4, 15))
    for i, (wl, s) in enumerate(zip(wls, sims)):
        s.synchronize()
        s.pull_from_gpu()
        lat_o, _ = run_oracle(wl, 2, 4, 7 + i, 25)
        np.testing.assert_array_equal(s.get_state().view(np.uint32), lat_o)


def test_batch_3d_per_phase():
    import gpucpm
    from gpucpm import workloads as W
    wls = [W.plain3d(L=L, ncells=n, seed=90 + i) for i, (L, n) in enumerate([(16, 5), (32, 40), (32, 60)])]
    print("CASE batch3d", flush=True)
    sims = []
    for i, wl in enumerate(wls):
        s = wl.make_cpm()
        s.push_to_gpu()
        s.set_seed(300 + i)
        sims.append(s)
    gpucpm.Batch(sims).run(**run_kwargs(2, 8, 6))
    for i, (wl, s) in enumerate(zip(wls, sims)):
        s.synchronize()
        s.pull_from_gpu()
        lat_o, tab_o = run_oracle(wl, 2, 8, 300 + i, 6)
        np.testing.assert_array_equal(s.get_state().view(np.uint32), lat_o)
        assert_tables_equal(s.get_cell_table(), tab_o, 3)


@pytest.mark.parametrize("c,u,shared", [(1, 1, 1), (2, 4, 0)])
def test_batch_cell_sync_equals_single_runs(c, u, shared):
    """cell_sync in a batch: each replica bit-identical to its own single run
    (which the single-run tests pin to the oracle)."""
    import gpucpm
    wls = _small_batch()[:4]
    print("CASE batch cell_sync", c, u, shared, flush=True)
    kw = run_kwargs(c, u, 6, cell_sync=1, shared=shared, positions_per_thread=2 * c)
    sims = []
    for i, wl in enumerate(wls):
        s = wl.make_cpm()
        s.push_to_gpu()
        s.set_seed(300 + i)
        sims.append(s)
    gpucpm.Batch(sims).run(**kw)
    for i, (wl, s) in enumerate(zip(wls, sims)):
        s.synchronize()
        s.pull_from_gpu()
        t = wl.make_cpm()
        t.push_to_gpu()
        t.set_seed(300 + i)
        t.run(**kw)
        t.synchronize()
        t.pull_from_gpu()
        np.testing.assert_array_equal(s.get_state(), t.get_state())


def test_batch_persistence_equals_single_runs():
    """3D replicas with a persistence term in a batch, chunked by
    inner_iterations: lattice and preferred directions bit-identical to
    single runs and to the oracle."""
    import gpucpm
    from gpucpm import workloads as W
    wls = [W.persistent(W.plain3d(L=32, ncells=30, seed=310 + i), lam=(0.0, 200.0 + 50 * i),
                        diffusion=(0.0, 0.6), history=2) for i in range(3)]
    print("CASE batch persistence", flush=True)
    kw = run_kwargs(1, 1, 6, inner=2)
    sims = []
    for i, wl in enumerate(wls):
        s = wl.make_cpm()
        s.push_to_gpu()
        s.set_seed(320 + i)
        sims.append(s)
    gpucpm.Batch(sims).run(**kw)
    for i, (wl, s) in enumerate(zip(wls, sims)):
        s.synchronize()
        s.pull_from_gpu()
        lat_o, tab_o = run_oracle(wl, 1, 1, 320 + i, 6, inner=2)
        np.testing.assert_array_equal(s.get_state().view(np.uint32), lat_o)
        assert_tables_equal(s.get_cell_table(), tab_o, 3)
        np.testing.assert_array_equal(s.get_pref_dirs().reshape(-1), tab_o.pref)


@pytest.mark.parametrize("form", [0, 1])
def test_batch_site_width_mixed(form):
    """Single runs on 16-bit sites, then a batch in which one replica's cells
    do not fit in 16 bits (so every replica runs on 32-bit sites), or the
    persistent batch form (32-bit sites): each replica continues its stream
    exactly."""
    import gpucpm
    from gpucpm import workloads as W
    wls = _small_batch()[:2] + [W.sorting2d(L=256, grid=128, edge=2, seed=41, ncells=16384)]
    print("CASE batch-width", form, [w.L for w in wls], flush=True)
    sims = []
    for i, wl in enumerate(wls):
        s = wl.make_cpm()
        s.push_to_gpu()
        s.set_seed(20 + i)
        s.run(**run_kwargs(1, 1, 3))
        sims.append(s)
    batch = gpucpm.Batch(sims)
    batch.set_form(form)
    batch.run(**run_kwargs(1, 1, 4))
    batch.synchronize()
    for i, (wl, s) in enumerate(zip(wls, sims)):
        s.pull_from_gpu()
        lat_o, tab_o = run_oracle(wl, 1, 1, 20 + i, 7)
        np.testing.assert_array_equal(s.get_state().view(np.uint32), lat_o)
        assert_tables_equal(s.get_cell_table(), tab_o, 2)


@pytest.mark.parametrize("att", ["1", "4", "8"])
def test_batch_attempts_per_thread(att):
    """The compacted batch kernel with 1, 4 or 8 attempts drawn per thread
    (each replica's block group covers its sub-blocks in chunks of att x 256)
    equals each replica's oracle run."""
    import gpucpm
    wls = _small_batch()
    print("CASE batch att", att, [w.L for w in wls], flush=True)
    sims = []
    for i, wl in enumerate(wls):
        s = wl.make_cpm()
        s.push_to_gpu()
        s.set_seed(300 + i)
        sims.append(s)
    batch = gpucpm.Batch(sims)
    batch.set_kernel_options(attempts=int(att))
    batch.run(**run_kwargs(1, 1, 12))
    for i, (wl, s) in enumerate(zip(wls, sims)):
        s.synchronize()
        s.pull_from_gpu()
        lat_o, tab_o = run_oracle(wl, 1, 1, 300 + i, 12)
        np.testing.assert_array_equal(s.get_state().view(np.uint32), lat_o)
        assert_tables_equal(s.get_cell_table(), tab_o, 2)


@pytest.mark.parametrize("cell_sync,rng", [(0, "philox"), (1, "philox"), (0, "minstd"),
                                           (1, "minstd")])
def test_batch_act_single_site_checkerboard(cell_sync, rng):
    """The act term with shared roles and positions_per_checkerboard = 1 in a
    batch: every phase is a decide launch and an apply launch for all
    replicas (with cell_sync the decide launch also claims).  Replicas with
    and without the term end bit-identical to their single runs and to the
    oracle, act lattices included; the persistent batch form refuses."""
    import gpucpm
    from gpucpm import workloads as W
    base = W.batched_sorting(nrep=4, seed=44, sizes=(32, 64))
    wls = [W.active(w) if i % 2 == 0 else w for i, w in enumerate(base)]
    print("CASE batch act c=1 cell_sync", cell_sync, rng, [w.L for w in wls], flush=True)
    kw = run_kwargs(1, 1, 16, 8, cell_sync=cell_sync)
    sims = []
    for i, wl in enumerate(wls):
        s = wl.make_cpm()
        s.push_to_gpu()
        s.set_seed(200 + i)
        s.set_rng(rng)
        sims.append(s)
    batch = gpucpm.Batch(sims)
    batch.run(**kw)
    batch.synchronize()
    for i, (wl, s) in enumerate(zip(wls, sims)):
        s.pull_from_gpu()
        lat = s.get_state().view(np.uint32).copy()
        t = wl.make_cpm()
        t.push_to_gpu()
        t.set_rng(rng)
        t.run(**kw, seed=200 + i)
        t.synchronize()
        t.pull_from_gpu()
        np.testing.assert_array_equal(lat, t.get_state().view(np.uint32))
        lat_o, tab_o = run_oracle(wl, 1, 1, 200 + i, 16, cell_sync=bool(cell_sync), rng=rng,
                                  inner=8 if rng == "minstd" else None)
        np.testing.assert_array_equal(lat, lat_o)
        assert_tables_equal(s.get_cell_table(), tab_o, 2)
        if i % 2 == 0:
            np.testing.assert_array_equal(s.get_act_state().view(np.uint32), tab_o.act)
        assert np.count_nonzero(lat != wl.lattice) > 0
    batch.set_form(1)
    with pytest.raises(NotImplementedError):
        batch.run(**kw)
