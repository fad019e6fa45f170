"""GPU parity of the reference-stream mode (Cpm.set_rng("minstd")): the
HIP sweep through the C ABI against the oracle's reference stream, which
tests/test_refstream.py pins against a thread-major restatement of the
reference kernel, including the reference's shared mode with
positions_per_thread > 2c, whose thread layout follows threads_per_block.
Every case prints its name before touching the GPU."""
from __future__ import annotations

import numpy as np
import pytest

from helpers import assert_tables_equal, run_kwargs, run_oracle, make_cpm

pytestmark = pytest.mark.gpu


def _announce(*a):
    print("CASE", *a, flush=True)


def _run(wl, c, u, seed, nmcs, inner, shared, ppt, cell_sync=0, tpb=4):
    sim = make_cpm(wl)
    sim.push_to_gpu()
    sim.set_seed(seed)
    sim.set_rng("minstd")
    assert sim.rng == "minstd"
    sim.run(**run_kwargs(c, u, nmcs, inner, shared=int(shared), positions_per_thread=ppt,
                         cell_sync=cell_sync, threads_per_block=tpb))
    sim.synchronize()
    sim.pull_from_gpu()
    return sim, sim.get_state().view(np.uint32).copy(), sim.get_cell_table()


def _cases():
    from gpucpm import workloads as W
    return [  # name, workload, c, u, shared, ppt, nmcs, inner
        ("2d-ns", W.sorting2d(L=128, grid=9, edge=10, seed=81), 2, 1, False, 4, 20, 5),
        ("2d-ns-u3", W.sorting2d(L=128, grid=9, edge=10, seed=82), 2, 3, False, 8, 12, 4),
        ("2d-sh", W.sorting2d(L=128, grid=9, edge=10, seed=83), 1, 1, True, 2, 20, 10),
        ("2d-sh-u4", W.sorting2d(L=64, grid=5, edge=10, seed=84), 2, 4, True, 4, 10, 10),
        ("2d-wall", W.walled(W.sorting2d(L=64, grid=5, edge=10, seed=85)), 2, 1, False, 4, 20, 7),
        ("2d-chem", W.chemotactic(W.sorting2d(L=64, grid=5, edge=10, seed=86)), 2, 1, False, 4, 15, 5),
        ("3d-ns", W.plain3d(L=32, ncells=60, seed=87), 1, 1, False, 2, 6, 3),
        ("3d-ns-u3", W.plain3d(L=32, ncells=60, seed=88), 2, 3, False, 4, 4, 2),
        ("3d-sh", W.plain3d(L=32, ncells=60, seed=89), 1, 1, True, 2, 6, 6),
        # shared roles, positions_per_thread > 2c: global blocks of
        # threads_per_block threads (the tpb of each case is in _TPB)
        ("2d-sh-gb", W.sorting2d(L=128, grid=9, edge=10, seed=94), 1, 1, True, 4, 12, 4),
        ("2d-sh-gb-u2", W.sorting2d(L=128, grid=9, edge=10, seed=95), 2, 2, True, 16, 8, 4),
        ("3d-sh-gb", W.plain3d(L=32, ncells=60, seed=96), 1, 1, True, 4, 4, 2),
    ]


_TPB = {"2d-sh-gb": 4, "2d-sh-gb-u2": 2, "3d-sh-gb": 2}


@pytest.mark.parametrize("narrow", ["1", "0"])
@pytest.mark.parametrize("idx", range(12))
def test_refstream_matches_oracle(idx, narrow, engine_options):
    engine_options(narrow=narrow)
    name, wl, c, u, shared, ppt, nmcs, inner = _cases()[idx]
    tpb = _TPB.get(name, 4)
    _announce("refstream", name, narrow, wl.L, wl.d, c, u, shared, ppt, tpb)
    sim, lat, table = _run(wl, c, u, 0x9E3779B9 + idx, nmcs, inner, shared, ppt, tpb=tpb)
    assert "k_sweep_phase_ref" in sim.last_launch()["kernel"]
    lat_o, tab_o = run_oracle(wl, c, u, 0x9E3779B9 + idx, nmcs, inner=inner, shared=shared, ppt=ppt,
                              rng="minstd", tpb=tpb)
    np.testing.assert_array_equal(lat, lat_o)
    assert_tables_equal(table, tab_o, wl.d)
    assert np.count_nonzero(lat != wl.lattice) > 0


def test_refstream_cell_sync_matches_oracle():
    from gpucpm import workloads as W
    wl = W.sorting2d(L=128, grid=9, edge=10, seed=90)
    _announce("refstream cell_sync", wl.L)
    _, lat, table = _run(wl, 2, 2, 31, 10, 5, False, 4, cell_sync=1)
    lat_o, tab_o = run_oracle(wl, 2, 2, 31, 10, inner=5, shared=False, ppt=4, rng="minstd",
                              cell_sync=True)
    np.testing.assert_array_equal(lat, lat_o)
    assert_tables_equal(table, tab_o, 2)


@pytest.mark.parametrize("d,shared", [(2, 0), (3, 0), (2, 1)])
def test_refstream_batch_equals_single_runs(d, shared):
    """A batch of reference-stream engines (each reseeded per chunk from its
    own seed) equals each engine's single run and the oracle."""
    import gpucpm
    from gpucpm import workloads as W
    if d == 2:
        wls = [W.walled(w) if i % 2 else w
               for i, w in enumerate(W.batched_sorting(nrep=4, seed=92, sizes=(32, 64, 128)))]
        c, u, ppt, n, inner = 2, 1, 4, 12, 4
        if shared:  # global blocks: two regions per thread and axis
            c, ppt = 1, 4
    else:
        wls = [W.plain3d(L=L, ncells=30, seed=93 + L) for L in (32, 64)]
        c, u, ppt, n, inner = 1, 1, 2, 4, 2
    _announce("refstream batch", d, shared, [w.L for w in wls])
    sims = []
    for i, wl in enumerate(wls):
        s = wl.make_cpm()
        s.push_to_gpu()
        s.set_seed(400 + i)
        s.set_rng("minstd")
        sims.append(s)
    gpucpm.Batch(sims).run(**run_kwargs(c, u, n, inner, shared=shared, positions_per_thread=ppt,
                                        threads_per_block=2))
    for i, (wl, s) in enumerate(zip(wls, sims)):
        s.synchronize()
        s.pull_from_gpu()
        lat_o, tab_o = run_oracle(wl, c, u, 400 + i, n, inner=inner, shared=bool(shared), ppt=ppt,
                                  rng="minstd", tpb=2)
        np.testing.assert_array_equal(s.get_state().view(np.uint32), lat_o)
        assert_tables_equal(s.get_cell_table(), tab_o, d)


def test_refstream_refusals_and_switching_back():
    """shared=1 with positions_per_thread > 2c where threads_per_block x
    positions_per_thread does not divide the lattice (no reference grid),
    batches mixing streams and the persistent batch form are refused; the
    counter stream is restored by set_rng("philox")."""
    import gpucpm
    from gpucpm import workloads as W
    wl = W.sorting2d(L=64, grid=5, edge=10, seed=91)
    _announce("refstream refusals", wl.L)
    sim = make_cpm(wl)
    sim.push_to_gpu()
    sim.set_rng("minstd")
    with pytest.raises(ValueError):  # 64 / 4 = 16 threads per axis: not blocks of 32
        sim.run(**run_kwargs(1, 1, 2, shared=1, positions_per_thread=4, threads_per_block=32))
    other = wl.make_cpm()
    other.push_to_gpu()
    with pytest.raises(ValueError):
        gpucpm.Batch([sim, other]).run(**run_kwargs(2, 1, 2, shared=0, positions_per_thread=4))
    b = gpucpm.Batch([sim])
    b.set_form(1)
    with pytest.raises(NotImplementedError):
        b.run(**run_kwargs(1, 1, 2, shared=1, positions_per_thread=2))
    with pytest.raises(ValueError):
        sim.set_rng("mt19937")
    sim.set_rng("philox")
    sim.set_seed(5)
    sim.run(**run_kwargs(1, 1, 5))
    sim.synchronize()
    sim.pull_from_gpu()
    lat_o, _ = run_oracle(wl, 1, 1, 5, 5)
    np.testing.assert_array_equal(sim.get_state().view(np.uint32), lat_o)
