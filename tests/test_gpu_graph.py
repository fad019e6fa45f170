"""Batched runs replayed from a hip graph (batch form 2): the per-phase
launches of G MCS are captured once, with the MCS number read from a device
counter the graph advances, and replayed.  Every replica must end
bit-identical to the oracle and to the per-phase batch; a graph is reused
while runs keep their shape and captured again when the shape changes; runs
the graph cannot express (cell_sync) fall back to per-phase launches; the
chosen form (-1) with the autotuner keeps the results exact while it samples
whole runs of each form."""
from __future__ import annotations

import numpy as np
import pytest

from helpers import assert_tables_equal, run_kwargs, run_oracle

pytestmark = pytest.mark.gpu


def _sims(wls, seed0, rng=None):
    sims = []
    for i, wl in enumerate(wls):
        s = wl.make_cpm()
        s.push_to_gpu()
        s.set_seed(seed0 + i)
        if rng:
            s.set_rng(rng)
        sims.append(s)
    return sims


def _lats(sims):
    out = []
    for s in sims:
        s.synchronize()
        s.pull_from_gpu()
        out.append(s.get_state().view(np.uint32).copy())
    return out


def _small_batch():
    from gpucpm import workloads as W
    return [W.walled(w) if i % 2 else (W.chemotactic(w) if i == 2 else w)
            for i, w in enumerate(W.batched_sorting(nrep=6, seed=40, sizes=(32, 64, 128)))]


def test_graph_replay_equals_oracle_and_reuses_the_graph():
    import gpucpm
    wls = _small_batch()
    print("CASE graph 2d", [w.L for w in wls], flush=True)
    sims = _sims(wls, 500)
    batch = gpucpm.Batch(sims)
    batch.set_form(2)
    batch.run(**run_kwargs(1, 1, 30, 10))      # G = 30: one capture, one replay
    assert batch.graph_info() == (2, 1, 1)
    lats = _lats(sims)
    for i, (wl, s, lat) in enumerate(zip(wls, sims, lats)):
        lat_o, tab_o = run_oracle(wl, 1, 1, 500 + i, 30)
        np.testing.assert_array_equal(lat, lat_o, err_msg=f"replica {i}")
        assert_tables_equal(s.get_cell_table(), tab_o, 2)
        assert s.mcs == 30
    batch.run(**run_kwargs(1, 1, 30, 10))      # the same shape: replayed again
    assert batch.graph_info() == (2, 1, 2)
    batch.run(**run_kwargs(1, 1, 128, 128))    # G = 64: a new capture, two replays
    assert batch.graph_info() == (2, 2, 4)
    lats = _lats(sims)
    # the same 188 MCS as per-phase batched launches
    ref = _sims(wls, 500)
    rb = gpucpm.Batch(ref)
    rb.set_form(0)
    for it, inner in ((30, 10), (30, 10), (128, 128)):
        rb.run(**run_kwargs(1, 1, it, inner))
    assert rb.graph_info()[0] == 0
    for i, (a, b) in enumerate(zip(lats, _lats(ref))):
        np.testing.assert_array_equal(a, b, err_msg=f"replica {i}")
    wl = wls[3]
    lat_o, tab_o = run_oracle(wl, 1, 1, 503, 188)
    np.testing.assert_array_equal(lats[3], lat_o)
    assert_tables_equal(sims[3].get_cell_table(), tab_o, 2)


@pytest.mark.parametrize("c,u,shared", [(2, 4, 0), (2, 2, 1)])
def test_graph_replay_colour_schedules(c, u, shared):
    """Several attempts per phase and both roles settings (uncompacted
    kernels), 16-bit sites."""
    import gpucpm
    wls = _small_batch()[:4]
    print("CASE graph schedule", c, u, shared, flush=True)
    kw = run_kwargs(c, u, 9, shared=shared, positions_per_thread=2 * c)
    sims = _sims(wls, 610)
    batch = gpucpm.Batch(sims)
    batch.set_form(2)
    batch.run(**kw)
    assert batch.graph_info()[0] == 2
    for i, (wl, s, lat) in enumerate(zip(wls, sims, _lats(sims))):
        lat_o, tab_o = run_oracle(wl, c, u, 610 + i, 9, shared=bool(shared))
        np.testing.assert_array_equal(lat, lat_o, err_msg=f"replica {i}")
        assert_tables_equal(s.get_cell_table(), tab_o, 2)


def test_graph_replay_3d():
    import gpucpm
    from gpucpm import workloads as W
    wls = [W.plain3d(L=L, ncells=n, seed=90 + i) for i, (L, n) in enumerate([(16, 5), (32, 40), (32, 60)])]
    print("CASE graph 3d", flush=True)
    sims = _sims(wls, 330)
    batch = gpucpm.Batch(sims)
    batch.set_form(2)
    batch.run(**run_kwargs(1, 1, 6))
    batch.run(**run_kwargs(1, 1, 6))
    assert batch.graph_info() == (2, 1, 2)
    for i, (wl, s, lat) in enumerate(zip(wls, sims, _lats(sims))):
        lat_o, tab_o = run_oracle(wl, 1, 1, 330 + i, 12)
        np.testing.assert_array_equal(lat, lat_o, err_msg=f"replica {i}")
        assert_tables_equal(s.get_cell_table(), tab_o, 3)


def test_graph_replay_act_split_phases():
    """The act term with shared roles and c = 1: each phase is a decide and
    an apply launch, both captured."""
    import gpucpm
    from gpucpm import workloads as W
    base = W.batched_sorting(nrep=4, seed=44, sizes=(32, 64))
    wls = [W.active(w) if i % 2 == 0 else w for i, w in enumerate(base)]
    print("CASE graph act", flush=True)
    kw = run_kwargs(1, 1, 16, 8)
    sims = _sims(wls, 200)
    batch = gpucpm.Batch(sims)
    batch.set_form(2)
    batch.run(**kw)
    assert batch.graph_info()[0] == 2
    for i, (wl, s, lat) in enumerate(zip(wls, sims, _lats(sims))):
        lat_o, tab_o = run_oracle(wl, 1, 1, 200 + i, 16)
        np.testing.assert_array_equal(lat, lat_o, err_msg=f"replica {i}")
        assert_tables_equal(s.get_cell_table(), tab_o, 2)
        if i % 2 == 0:
            np.testing.assert_array_equal(s.get_act_state().view(np.uint32), tab_o.act)


def test_graph_form_falls_back_where_it_cannot_apply():
    """cell_sync grows its phase tags per launch: the graph form runs the
    per-phase launches instead, with the same results as form 0."""
    import gpucpm
    wls = _small_batch()[:3]
    print("CASE graph fallback", flush=True)
    kw = run_kwargs(1, 1, 6, cell_sync=1)
    sims = _sims(wls, 300)
    batch = gpucpm.Batch(sims)
    batch.set_form(2)
    batch.run(**kw)
    assert batch.graph_info() == (0, 0, 0)
    ref = _sims(wls, 300)
    gpucpm.Batch(ref).run(**kw)
    for i, (a, b) in enumerate(zip(_lats(sims), _lats(ref))):
        np.testing.assert_array_equal(a, b, err_msg=f"replica {i}")


def test_chosen_form_with_autotune_stays_exact():
    """Form -1 with the autotuner: whole runs alternate between per-phase
    launches, the persistent launch and graph replay until each has its
    samples; every run continues each replica's stream exactly."""
    import gpucpm
    wls = _small_batch()[:4]
    print("CASE graph autotune", flush=True)
    sims = _sims(wls, 900)
    batch = gpucpm.Batch(sims)
    batch.set_form(-1)
    batch.set_autotune(True)
    used = set()
    nrun = 30
    for _ in range(nrun):
        batch.run(**run_kwargs(1, 1, 4, 2))
        used.add(batch.graph_info()[0])
    assert used >= {0, 1, 2}
    for i, (wl, s, lat) in enumerate(zip(wls, sims, _lats(sims))):
        lat_o, tab_o = run_oracle(wl, 1, 1, 900 + i, 4 * nrun, inner=2)
        np.testing.assert_array_equal(lat, lat_o, err_msg=f"replica {i}")
        assert_tables_equal(s.get_cell_table(), tab_o, 2)
